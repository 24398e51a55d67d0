"""Host-side headers of a compiled mechanism.

For each mechanism the reference's generator renders a C header with the
mechanism's metadata macros (src/mkpp/templates/host_api/mkpp.h.j2:17-72)
and a C++ header whose `SolverKernels<Device>` type carries the species
enumeration, the runtime-parameter and environment flags and the
generation-time tolerances and step seed (src/mkpp/templates/header.j2:65-269).
This module writes both for mkpp_amd:

* ``<mech>/mkpp.h``: the reference's C declarations and return codes with
  MKPP_NUM_RATE_PARAMETERS / MKPP_NUM_SPECIES / MKPP_NUM_REACTIONS /
  MKPP_NUM_PHOTOLYSIS, plus the drop-in's generation-time solver defaults;
* ``<mech>.hpp``: ``mkpp::generated::<mech>::SolverKernels<MemorySpace>``,
  whose ``integrate`` runs every cell of a (cells x species) view through the
  C ABI with the type's solver options;
* ``<mech>/mkpp_mod.f90``: the Fortran module of the drop-in
  (src/mkpp/templates/host_api/mkpp_mod.f90.j2:54-169);
* ``mkpp_host/mechanisms.hpp``: the includes of every bundled mechanism's
  header that is on the include path, and the lookup of its generation-time
  solver options by name (what ``MechanismRegistry`` dispatches on).

The solver defaults come from ``--solver`` and the environment file's
``solver:`` block, as the reference's ``mkpp compile`` takes them
(src/mkpp/cli.py:229-236, 394-399; src/mkpp/parser.py:370-434); without
either they are the reference's: ROS-3, atol 1e-3, rtol 1e-6, seed 1e-6
(src/mkpp/template_context.py:283-317).
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

from .tableau import METHOD_IDS

METHOD_MACROS = {"ros2": "MKPP_METHOD_ROS2", "ros3": "MKPP_METHOD_ROS3", "ros4": "MKPP_METHOD_ROS4",
                 "rodas3": "MKPP_METHOD_RODAS3", "rodas4": "MKPP_METHOD_RODAS4"}


@dataclass
class SolverDefaults:
    """Generation-time solver configuration of a drop-in."""

    method: str = "ros3"
    atol: list = field(default_factory=list)  # per species (physical order); empty: 1e-3 each
    rtol: list = field(default_factory=list)  # empty: 1e-6 each
    initial_step_seed: float = 1.0e-6

    def per_species(self, n: int) -> tuple[list, list]:
        a = list(self.atol) if self.atol else [1.0e-3] * n
        r = list(self.rtol) if self.rtol else [1.0e-6] * n
        if len(a) != n or len(r) != n:
            raise ValueError(f"tolerance arrays must have {n} entries")
        return [float(x) for x in a], [float(x) for x in r]


def load_environment(path: str) -> dict:
    """The `solver:` block (atol, rtol, initial_step_seed) and the environment
    block of an environment file, read as the reference reads it
    (src/mkpp/parser.py:370-434): JSON or YAML; the environment under
    `environment`, `meteorology` or the top level."""
    import json

    import yaml

    with open(path, encoding="utf-8") as f:
        text = f.read()
    data = json.loads(text) if path.endswith(".json") else yaml.safe_load(text)
    if not isinstance(data, dict):
        raise ValueError(f"environment file {path} must hold a key-value mapping")
    env = data.get("environment", data.get("meteorology", data))
    solver = data.get("solver", {})
    if not isinstance(solver, dict):
        solver = {}
    return {
        "temperature": float(env.get("temperature", env.get("T", 298.15))),
        "pressure": float(env.get("pressure", env.get("P", 101325.0))),
        "air_density": float(env.get("air_density", env.get("M", 2.46e19))),
        "relative_humidity": float(env.get("relative_humidity", env.get("RH", 0.5))),
        "atol": None if solver.get("atol") is None else float(solver["atol"]),
        "rtol": None if solver.get("rtol") is None else float(solver["rtol"]),
        "initial_step_seed": None if solver.get("initial_step_seed") is None else float(solver["initial_step_seed"]),
    }


def defaults_from(method: str = "ros3", env_path: str | None = None, num_species: int = 0) -> SolverDefaults:
    if method not in METHOD_IDS:
        raise ValueError(f"unknown solver {method!r}; choose one of {', '.join(METHOD_IDS)}")
    d = SolverDefaults(method=method)
    if env_path:
        env = load_environment(env_path)
        if env["atol"] is not None:
            d.atol = [env["atol"]] * num_species
        if env["rtol"] is not None:
            d.rtol = [env["rtol"]] * num_species
        if env["initial_step_seed"] is not None:
            d.initial_step_seed = env["initial_step_seed"]
    return d


def _num(x: float) -> str:
    return repr(float(x))


def _meta(p: dict) -> dict:
    n = len(p["species"])
    P = int(p["num_runtime_parameters"])
    return {"name": p["name"], "n": n, "P": P, "reactions": int(p["num_reactions"]),
            "photolysis": P, "runtime_env": bool(p["supports_runtime_environment"]), "species": list(p["species"])}


def c_header(p: dict, d: SolverDefaults) -> str:
    """Self-contained per-mechanism mkpp.h (mkpp.h.j2:17-72 plus defaults)."""
    m = _meta(p)
    at, rt = d.per_species(m["n"])
    uniform = len(set(at)) == 1 and len(set(rt)) == 1
    guard = f"MKPP_AMD_GENERATED_{m['name'].upper()}_H"
    L = [
        "/*",
        f" * mkpp.h -- C ABI of the compiled mechanism {m['name']} (libmkpp_{m['name']}.so).",
        " * Generated by mkpp_amd.gen.emit_host; do not edit.",
        " *",
        " * The reference's generated C header (src/mkpp/templates/host_api/mkpp.h.j2:17-72):",
        " * the same symbols, argument meaning, LayoutLeft buffers and return codes,",
        " * and the mechanism's metadata macros.  The MKPP_DEFAULT_* macros state the",
        " * generation-time solver options this drop-in's handles start with.",
        " */",
        f"#ifndef {guard}",
        f"#define {guard}",
        "",
        "#include <stddef.h>",
        "",
        "#ifdef __cplusplus",
        'extern "C" {',
        "#endif",
        "",
        "#define MKPP_SUCCESS 0",
        "#define MKPP_ERROR_NULL_POINTER -101",
        "#define MKPP_ERROR_INVALID_HANDLE -102",
        "#define MKPP_ERROR_INVALID_ARGUMENT -103",
        "#define MKPP_ERROR_INVALID_STATE -104",
        "#define MKPP_ERROR_CONVERGENCE -105",
        "#define MKPP_ERROR_NOT_FOUND -106",
        "",
        "typedef struct mkpp_handle_s* mkpp_handle_t;",
        "",
        "/* Mechanism metadata (mkpp.h.j2:30-33). */",
        f"#define MKPP_NUM_RATE_PARAMETERS {m['P']}",
        f"#define MKPP_NUM_SPECIES {m['n']}",
        f"#define MKPP_NUM_REACTIONS {m['reactions']}",
        f"#define MKPP_NUM_PHOTOLYSIS {m['photolysis']}",
        "",
        "/* Generation-time solver options of this drop-in: method id (0 ROS-2,",
        " * 1 ROS-3, 2 ROS-4, 3 RODAS-3, 4 RODAS-4), tolerances and the first step as",
        " * a fraction of the chemistry interval. */",
        f"#define MKPP_DEFAULT_METHOD {METHOD_IDS[d.method]}",
    ]
    if uniform:
        L += [f"#define MKPP_DEFAULT_ATOL {_num(at[0])}", f"#define MKPP_DEFAULT_RTOL {_num(rt[0])}"]
    L += [
        f"#define MKPP_DEFAULT_INITIAL_STEP_SEED {_num(d.initial_step_seed)}",
        "",
        "int mkpp_create_handle(mkpp_handle_t* handle_out, size_t num_cells);",
        "int mkpp_destroy_handle(mkpp_handle_t handle);",
        "int mkpp_initialize(mkpp_handle_t handle);",
        "int mkpp_get_species_count(mkpp_handle_t handle, size_t* count_out);",
        "int mkpp_get_reaction_count(mkpp_handle_t handle, size_t* count_out);",
        "int mkpp_get_photolysis_count(mkpp_handle_t handle, size_t* count_out);",
        "int mkpp_get_species_index(mkpp_handle_t handle, const char* species_name, int* index_out);",
        "int mkpp_get_photolysis_index(mkpp_handle_t handle, const char* photo_name, int* index_out);",
        "int mkpp_set_state_ptrs(mkpp_handle_t handle, double* species_conc, const double* temp, const double* pressure,",
        "                        const double* air_density);",
        "/* LayoutLeft: values[cell + num_cells * parameter], MKPP_NUM_RATE_PARAMETERS parameters per cell. */",
        "int mkpp_set_rate_parameter_ptrs(mkpp_handle_t handle, const double* values, size_t count);",
        "int mkpp_set_photolysis_ptrs(mkpp_handle_t handle, const double* photolysis_rates);",
        "int mkpp_set_cloud_water_ptr(mkpp_handle_t handle, const double* values);",
        "int mkpp_set_aerosol_ptrs(mkpp_handle_t handle, const double* relative_humidity, const double* aerosol_surface_area);",
        "int mkpp_integrate(mkpp_handle_t handle, double dt);",
        "/* rates_out[cell + num_cells * species]; jac_out[cell + num_cells * (i + MKPP_NUM_SPECIES * j)] = dF_i/dC_j. */",
        "int mkpp_evaluate_rates(mkpp_handle_t handle, double* rates_out);",
        "int mkpp_evaluate_jacobian(mkpp_handle_t handle, double* jac_out);",
        "int mkpp_get_last_error(mkpp_handle_t handle, char* buffer, size_t buffer_len);",
        "",
        "#ifdef __cplusplus",
        "}",
        "#endif",
        "",
        f"#endif /* {guard} */",
    ]
    return "\n".join(L) + "\n"


def cpp_header(p: dict, d: SolverDefaults) -> str:
    """<mech>.hpp: mkpp::generated::<mech>::SolverKernels (header.j2:65-269)."""
    m = _meta(p)
    at, rt = d.per_species(m["n"])
    name = m["name"]

    def arr(vals):
        out, line = [], "      "
        for i, v in enumerate(vals):
            item = _num(v) + (", " if i + 1 < len(vals) else "")
            if len(line) + len(item) > 110:
                out.append(line.rstrip())
                line = "      "
            line += item
        out.append(line.rstrip())
        return "\n".join(out)

    L = [
        f"// {name}.hpp -- SolverKernels of the compiled mechanism {name}.",
        "// Generated by mkpp_amd.gen.emit_host; do not edit.",
        "//",
        "// The reference's generated mechanism header (src/mkpp/templates/header.j2:65-269)",
        "// declares the Species enumeration and SolverKernels<DeviceType> with the",
        "// mechanism's constants and a per-cell integrate() run inside a Kokkos",
        "// parallel_for.  Here one integrate() call runs every cell of a (cells x",
        "// species) view on the GPU through the C ABI of libmkpp_amd.so, with the",
        "// solver options of this type: the generation-time method and tolerances",
        "// below and the initial_step_seed member (which a caller may override).",
        "#pragma once",
        '#include "mkpp_host/dispatcher.hpp"',
        "",
        f"namespace mkpp::generated::{name} {{",
        "",
        "// Species indices into the state (physical order).",
        "enum Species : int {",
    ]
    for i, s in enumerate(m["species"]):
        L.append(f"  {s} = {i}{',' if i + 1 < m['n'] else ''}")
    L += [
        "};",
        "",
        "template <class DeviceType = ::mkpp::host::DefaultMemorySpace>",
        "struct SolverKernels {",
        "  using memory_space = DeviceType;",
        f'  static constexpr const char* NAME = "{name}";',
        f"  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = {'true' if m['runtime_env'] else 'false'};",
        f"  static constexpr int NUM_RUNTIME_PARAMETERS = {m['P']};",
        "  static constexpr bool HAS_CLOUD_GATED = false;",
        f"  static constexpr int NUM_SPECIES = {m['n']};",
        f"  static constexpr int NUM_REACTIONS = {m['reactions']};",
        f"  static constexpr int NUM_PHOTOLYSIS = {m['photolysis']};",
        f"  // generation-time solver ({d.method}) and tolerances, physical species order",
        f"  static constexpr int METHOD = {METHOD_IDS[d.method]};",
        "  static constexpr double atol[NUM_SPECIES] = {",
        arr(at),
        "  };",
        "  static constexpr double rtol[NUM_SPECIES] = {",
        arr(rt),
        "  };",
        f"  double initial_step_seed = {_num(d.initial_step_seed)};",
        "",
        "  // Integrates every cell of `state` (extent(0) cells x NUM_SPECIES, cells",
        "  // contiguous for device memory) over params.dt, params.steps times, with",
        "  // the per-cell overrides of `params` (dispatcher.hpp:197-255).",
        "  template <class StateView, class MemorySpace>",
        "  void integrate(StateView state, const ::mkpp::host::BasicHostExecutionParams<MemorySpace>& params) const {",
        "    ::mkpp::host::SolverOptions o;",
        "    o.method = METHOD;",
        "    o.atol = atol;",
        "    o.rtol = rtol;",
        "    o.initial_step_seed = initial_step_seed;",
        "    ::mkpp::host::detail::execute_with(NAME, NUM_SPECIES, state, params, o);",
        "  }",
        "};",
        "",
        f"}}  // namespace mkpp::generated::{name}",
        "",
        f"namespace mech_{name} {{",
        f"namespace mkpp = ::mkpp::generated::{name};",
        "}",
    ]
    return "\n".join(L) + "\n"


def mechanisms_header(names: list) -> str:
    """mkpp_host/mechanisms.hpp: the bundled mechanisms' headers that are on the
    include path, and their solver options by name (mechanisms.hpp:5-43)."""
    L = [
        "// mechanisms.hpp -- the compiled mechanisms' SolverKernels headers.",
        "// Generated by mkpp_amd.gen.emit_host; do not edit.",
        "//",
        "// As in the reference (src/mkpp_host/include/mkpp_host/mechanisms.hpp:5-43),",
        "// each mechanism's <name>.hpp is included when it is on the include path;",
        "// MechanismRegistry then runs that mechanism with the header's",
        "// generation-time solver options, and every other mechanism of",
        "// libmkpp_amd.so with the library's defaults.",
        "#pragma once",
        "#include <string>",
        "",
        '#include "dispatcher.hpp"',
        "",
    ]
    for nm in names:
        L += [f'#if __has_include("{nm}.hpp")', f'#include "{nm}.hpp"', f"#define MKPP_HOST_HAS_{nm.upper()} 1",
              "#endif"]
    L += [
        "",
        "namespace mkpp::host::detail {",
        "",
        "// The generation-time solver options of `name` if its header is included.",
        "inline bool generated_options(const std::string& name, SolverOptions& o) {",
    ]
    for nm in names:
        L += [
            f"#ifdef MKPP_HOST_HAS_{nm.upper()}",
            f'  if (name == "{nm}") {{',
            f"    using K = ::mkpp::generated::{nm}::SolverKernels<>;",
            "    o.method = K::METHOD;",
            "    o.atol = K::atol;",
            "    o.rtol = K::rtol;",
            "    o.initial_step_seed = K{}.initial_step_seed;",
            "    return true;",
            "  }",
            "#endif",
        ]
    L += ["  (void)name;", "  (void)o;", "  return false;", "}", "", "}  // namespace mkpp::host::detail"]
    return "\n".join(L) + "\n"


def fortran_module(p: dict, d: SolverDefaults) -> str:
    """<mech>/mkpp_mod.f90: the Fortran module of the drop-in C ABI with the
    reference's procedure names and argument meaning
    (src/mkpp/templates/host_api/mkpp_mod.f90.j2:54-169): each procedure
    returns its C return code in `status`; species and photolysis indices are
    1-based; arrays are (cells, species) and (cells, parameters), which is the
    C ABI's LayoutLeft order."""
    m = _meta(p)
    L = [
        "!" + "=" * 78,
        f"! mkpp_mod.f90 -- Fortran binding of the compiled mechanism {m['name']} (libmkpp_{m['name']}.so).",
        "! Generated by mkpp_amd.gen.emit_host; do not edit.",
        "!" + "=" * 78,
        "module mkpp_mod",
        "  use, intrinsic :: iso_c_binding",
        "  implicit none",
        "  private",
        "",
        "  integer(c_int), parameter, public :: MKPP_SUCCESS = 0_c_int",
        "  integer(c_int), parameter, public :: MKPP_ERROR_NULL_PTR = -101_c_int",
        "  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_HDL = -102_c_int",
        "  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_ARG = -103_c_int",
        "  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_STE = -104_c_int",
        "  integer(c_int), parameter, public :: MKPP_ERROR_CONVERGENCE = -105_c_int",
        "  integer(c_int), parameter, public :: MKPP_ERROR_NOT_FOUND = -106_c_int",
        "",
        f"  integer, parameter, public :: MKPP_NUM_RATE_PARAMETERS = {m['P']}",
        f"  integer, parameter, public :: MKPP_NUM_SPECIES = {m['n']}",
        f"  integer, parameter, public :: MKPP_NUM_REACTIONS = {m['reactions']}",
        f"  integer, parameter, public :: MKPP_NUM_PHOTOLYSIS = {m['photolysis']}",
        "  ! generation-time solver options of this drop-in (method id 0 ROS-2 ... 4 RODAS-4)",
        f"  integer, parameter, public :: MKPP_DEFAULT_METHOD = {METHOD_IDS[d.method]}",
        "",
        "  type, public :: mkpp_handle_t",
        "    type(c_ptr) :: ptr = c_null_ptr",
        "  end type mkpp_handle_t",
        "",
    ]
    procs = ["create_handle", "destroy_handle", "initialize", "get_species_count", "get_reaction_count",
             "get_photolysis_count", "get_species_index", "get_photolysis_index", "set_state_ptrs",
             "set_rate_parameter_ptrs", "set_photolysis_ptrs", "set_cloud_water_ptr", "set_aerosol_ptrs", "integrate",
             "evaluate_rates", "evaluate_jacobian", "get_last_error"]
    for pr in procs:
        L.append(f"  public :: mkpp_{pr}")
    # C interfaces: (name, [(arg, declaration)])
    h = ("handle", "type(c_ptr), value :: handle")
    cints = {
        "create_handle": [("handle_out", "type(c_ptr), intent(out) :: handle_out"),
                          ("num_cells", "integer(c_size_t), value :: num_cells")],
        "destroy_handle": [h],
        "initialize": [h],
        "get_species_count": [h, ("count_out", "integer(c_size_t), intent(out) :: count_out")],
        "get_reaction_count": [h, ("count_out", "integer(c_size_t), intent(out) :: count_out")],
        "get_photolysis_count": [h, ("count_out", "integer(c_size_t), intent(out) :: count_out")],
        "get_species_index": [h, ("name", "character(kind=c_char), intent(in) :: name(*)"),
                              ("index_out", "integer(c_int), intent(out) :: index_out")],
        "get_photolysis_index": [h, ("name", "character(kind=c_char), intent(in) :: name(*)"),
                                 ("index_out", "integer(c_int), intent(out) :: index_out")],
        "set_state_ptrs": [h] + [(a, f"type(c_ptr), value :: {a}") for a in ("conc", "temp", "pres", "rho")],
        "set_rate_parameter_ptrs": [h, ("values", "type(c_ptr), value :: values"),
                                    ("count", "integer(c_size_t), value :: count")],
        "set_photolysis_ptrs": [h, ("values", "type(c_ptr), value :: values")],
        "set_cloud_water_ptr": [h, ("values", "type(c_ptr), value :: values")],
        "set_aerosol_ptrs": [h, ("rh", "type(c_ptr), value :: rh"), ("sa", "type(c_ptr), value :: sa")],
        "integrate": [h, ("dt", "real(c_double), value :: dt")],
        "evaluate_rates": [h, ("out", "type(c_ptr), value :: out")],
        "evaluate_jacobian": [h, ("out", "type(c_ptr), value :: out")],
        "get_last_error": [h, ("buffer", "character(kind=c_char), intent(out) :: buffer(*)"),
                           ("buffer_len", "integer(c_size_t), value :: buffer_len")],
    }
    L += ["", "  interface"]
    for pr in procs:
        args = cints[pr]
        L.append(f"    integer(c_int) function c_{pr}({', '.join(a for a, _ in args)}) bind(C, name=\"mkpp_{pr}\")")
        L.append("      import :: c_ptr, c_int, c_size_t, c_double, c_char")
        for _, decl in args:
            L.append(f"      {decl}")
        L.append(f"    end function c_{pr}")
    L += ["  end interface", "", "contains", ""]
    hd = "    type(mkpp_handle_t), intent(in) :: handle"
    st = "    integer, intent(out) :: status"
    body = {
        "create_handle": (["handle", "num_cells", "status"],
                          ["    type(mkpp_handle_t), intent(out) :: handle", "    integer, intent(in) :: num_cells", st],
                          ["    status = int(c_create_handle(handle%ptr, int(num_cells, c_size_t)))"]),
        "destroy_handle": (["handle", "status"], ["    type(mkpp_handle_t), intent(inout) :: handle", st],
                           ["    status = int(c_destroy_handle(handle%ptr))", "    handle%ptr = c_null_ptr"]),
        "initialize": (["handle", "status"], [hd, st], ["    status = int(c_initialize(handle%ptr))"]),
    }
    for cnt in ("species", "reaction", "photolysis"):
        body[f"get_{cnt}_count"] = (["handle", "count_out", "status"],
                                    [hd, "    integer, intent(out) :: count_out", st, "    integer(c_size_t) :: n"],
                                    [f"    status = int(c_get_{cnt}_count(handle%ptr, n))", "    count_out = int(n)"])
    for what in ("species", "photolysis"):
        body[f"get_{what}_index"] = (
            ["handle", "name", "index_out", "status"],
            [hd, "    character(len=*), intent(in) :: name", "    integer, intent(out) :: index_out", st,
             "    integer(c_int) :: i0"],
            [f"    status = int(c_get_{what}_index(handle%ptr, trim(name) // c_null_char, i0))",
             "    index_out = -1",
             "    if (status == MKPP_SUCCESS) index_out = int(i0) + 1  ! 1-based"])
    body["set_state_ptrs"] = (
        ["handle", "conc", "temp", "pres", "rho", "status"],
        [hd, "    real(c_double), target, intent(inout) :: conc(:, :)  ! (cells, species)",
         "    real(c_double), target, intent(in) :: temp(:), pres(:), rho(:)", st],
        ["    if (.not. (is_contiguous(conc) .and. is_contiguous(temp) .and. is_contiguous(pres) .and. &",
         "               is_contiguous(rho))) then",
         "      status = MKPP_ERROR_INVALID_ARG",
         "      return",
         "    end if",
         "    status = int(c_set_state_ptrs(handle%ptr, c_loc(conc), c_loc(temp), c_loc(pres), c_loc(rho)))"])
    body["set_rate_parameter_ptrs"] = (
        ["handle", "values", "status"],
        [hd, "    real(c_double), target, intent(in) :: values(:, :)  ! (cells, MKPP_NUM_RATE_PARAMETERS)", st],
        ["    if (.not. is_contiguous(values) .or. size(values, 2) /= MKPP_NUM_RATE_PARAMETERS) then",
         "      status = MKPP_ERROR_INVALID_ARG",
         "      return",
         "    end if",
         "    status = int(c_set_rate_parameter_ptrs(handle%ptr, c_loc(values), int(size(values, 2), c_size_t)))"])
    body["set_photolysis_ptrs"] = (
        ["handle", "values", "status"],
        [hd, "    real(c_double), target, intent(in) :: values(:, :)  ! (cells, MKPP_NUM_PHOTOLYSIS)", st],
        ["    if (.not. is_contiguous(values)) then",
         "      status = MKPP_ERROR_INVALID_ARG",
         "      return",
         "    end if",
         "    status = int(c_set_photolysis_ptrs(handle%ptr, c_loc(values)))"])
    body["set_cloud_water_ptr"] = (
        ["handle", "values", "status"], [hd, "    real(c_double), target, intent(in) :: values(:)", st],
        ["    status = int(c_set_cloud_water_ptr(handle%ptr, c_loc(values)))"])
    body["set_aerosol_ptrs"] = (
        ["handle", "rh", "sa", "status"], [hd, "    real(c_double), target, intent(in) :: rh(:), sa(:)", st],
        ["    status = int(c_set_aerosol_ptrs(handle%ptr, c_loc(rh), c_loc(sa)))"])
    body["integrate"] = (["handle", "dt", "status"], [hd, "    real(c_double), intent(in) :: dt", st],
                         ["    status = int(c_integrate(handle%ptr, dt))"])
    body["evaluate_rates"] = (
        ["handle", "rates", "status"], [hd, "    real(c_double), target, intent(out) :: rates(:, :)  ! (cells, species)", st],
        ["    status = int(c_evaluate_rates(handle%ptr, c_loc(rates)))"])
    body["evaluate_jacobian"] = (
        ["handle", "jac", "status"],
        [hd, "    real(c_double), target, intent(out) :: jac(:, :)  ! (cells, species * species), column j*N + i", st],
        ["    status = int(c_evaluate_jacobian(handle%ptr, c_loc(jac)))"])
    body["get_last_error"] = (
        ["handle", "message"],
        [hd, "    character(len=*), intent(out) :: message", "    character(kind=c_char) :: buf(512)",
         "    integer(c_int) :: rc", "    integer :: i"],
        ["    buf = c_null_char", "    rc = c_get_last_error(handle%ptr, buf, 512_c_size_t)", "    message = ''",
         "    do i = 1, min(len(message), 512)", "      if (buf(i) == c_null_char) exit",
         "      message(i:i) = buf(i)", "    end do"])
    for pr in procs:
        args, decls, stmts = body[pr]
        L.append(f"  subroutine mkpp_{pr}({', '.join(args)})")
        L += decls
        L += stmts
        L.append(f"  end subroutine mkpp_{pr}")
        L.append("")
    L.append("end module mkpp_mod")
    return "\n".join(L) + "\n"


def write_mechanism(p: dict, d: SolverDefaults, include_dir: str, c_dir: str | None = None) -> list:
    """Writes <include_dir>/<mech>.hpp and <c_dir or include_dir/<mech>>/mkpp.h."""
    name = p["name"]
    c_dir = c_dir or os.path.join(include_dir, name)
    os.makedirs(include_dir, exist_ok=True)
    os.makedirs(c_dir, exist_ok=True)
    out = []
    for path, text in ((os.path.join(include_dir, f"{name}.hpp"), cpp_header(p, d)),
                       (os.path.join(c_dir, "mkpp.h"), c_header(p, d)),
                       (os.path.join(c_dir, "mkpp_mod.f90"), fortran_module(p, d))):
        with open(path, "w", encoding="utf-8") as f:
            f.write(text)
        out.append(path)
    return out


def write_mechanisms_list(names: list, host_dir: str) -> str:
    os.makedirs(host_dir, exist_ok=True)
    path = os.path.join(host_dir, "mechanisms.hpp")
    with open(path, "w", encoding="utf-8") as f:
        f.write(mechanisms_header(names))
    return path
