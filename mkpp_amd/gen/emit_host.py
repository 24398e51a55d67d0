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


def write_mechanism(p: dict, d: SolverDefaults, include_dir: str, c_dir: str | None = None) -> list:
    """Writes <include_dir>/<mech>.hpp and <c_dir or include_dir/<mech>>/mkpp.h."""
    name = p["name"]
    c_dir = c_dir or os.path.join(include_dir, name)
    os.makedirs(include_dir, exist_ok=True)
    os.makedirs(c_dir, exist_ok=True)
    out = []
    for path, text in ((os.path.join(include_dir, f"{name}.hpp"), cpp_header(p, d)),
                       (os.path.join(c_dir, "mkpp.h"), c_header(p, d))):
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
