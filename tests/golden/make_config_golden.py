"""Writes tests/golden/config_cases.json by running the REFERENCE's own
configuration parsers (oracle/_ref/ref_config, built from the reference tree
by oracle/build_ref.sh: parse_solver_config and parse_boundary_config with the
vendored toml++ 3.4.0) on configuration files written here.

Every case holds the file's text (written here, not taken from the reference)
and either the parsed structure the reference produced or the message its
mpi::fatal printed; boundary-condition cases also hold the log_stat warnings.
The parsed structures of the reference's own example files
(examples/*/*.toml) are stored too, under "ref_examples" (values only; the
test reads those files where the reference tree is present).

usage: python tests/golden/make_config_golden.py
"""
import glob
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("CFD_REFERENCE", "/root/reference")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_config")

BASE = {
    "flow": 'eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = 1.4\ngas_constant = 287.052874\n',
    "initial": "rho = 1.1768\npressure = 101325.0\nvelocity = [1041.0, 0.0, 0.0]\n",
    "numerics": 'flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\n',
    "time": 'mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 25.0\nmax_iterations = 500\nresidual_tolerance = 1.0e-10\n',
    "output": 'directory = "out/case"\nfield_interval = 100\nresidual_interval = 10\n',
}


def solver_text(drop=(), **over):
    """A solver file: the BASE sections, with sections replaced (string body)
    or removed (drop), plus extra sections in `over`; `_raw*` entries are
    written first, at the root level."""
    parts = [body for name, body in over.items() if name.startswith("_raw")]  # root-level lines come first
    secs = dict(BASE)
    secs.update({k: v for k, v in over.items() if not k.startswith("_raw")})
    for name, body in secs.items():
        if name in drop or body is None:
            continue
        parts.append(f"[{name}]\n{body}")
    return "\n".join(parts)


def solver_cases():
    S = solver_text
    visc = 'eos = "IDEAL_GAS"\nflow_model = "VISCOUS_FLOW"\nprandtl = 0.72\ngamma = 1.4\ngas_constant = 287.0\n'
    c = {
        # valid files
        "minimal_first_order": S(numerics='flux = "HLLC"\nreconstruction = "FIRST_ORDER"\n'),
        "muscl_venkat_defaults": S(),
        "viscous_barth_gg_cell": S(flow=visc, numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "BARTH"\n'
                                   'venkat_k = 2.5\ngradient = "GREEN_GAUSS_CELL"\n'),
        "van_albada_lsq_face": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VAN_ALBADA"\n'
                                 'gradient = "LEAST_SQUARES_FACE"\n'),
        "lsq_node": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\ngradient = "LEAST_SQUARES_NODE"\n'),
        "directional_minmod": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL_DIRECTIONAL"\nlimiter = "MINMOD_1D"\n'),
        "directional_van_albada": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL_DIRECTIONAL"\nlimiter = "VAN_ALBADA_1D"\n'),
        "first_order_with_limiter_key": S(numerics='flux = "HLLC"\nreconstruction = "FIRST_ORDER"\nlimiter = "NONSENSE"\n'),
        "unsteady_bdf1": S(time='mode = "Unsteady"\nscheme = "BACKWARD_EULER"\ncfl = 5\nmax_iterations = 30\n'
                           'residual_tolerance = 1e-3\ntime_step = 2.5e-4\nmax_time_steps = 40\nbdf_order = 1\n'),
        "unsteady_default_bdf": S(time='mode = "Unsteady"\nscheme = "FORWARD_EULER"\ncfl = 0.5\nmax_iterations = 30\n'
                                  'residual_tolerance = 1e-3\ntime_step = 1e-5\nmax_time_steps = 3\n'),
        "unknown_mode_stays_steady": S(time='mode = "Transient"\nscheme = "BACKWARD_EULER"\ncfl = 5.0\nmax_iterations = 3\n'
                                       'residual_tolerance = 1e-3\ntime_step = 1.0\n'),
        "turbulence_defaults": S(flow=visc, turbulence='model = "SA"\n'),
        "turbulence_full": S(flow=visc, turbulence='model = "SA"\nnu_inf_ratio = 5\nmax_distance_sweeps = 200.0\n'
                             'distance_tolerance = 1e-6\n'),
        "linalg_defaults": S(linalg=""),
        "linalg_full": S(linalg='solver = "BICGSTAB"\npreconditioner = "SGS"\nrel_tol = 1e-2\nabs_tol = 1e-14\n'
                         'max_iter = 250\nverbosity = "Verbose"\nres_verify = true\n'),
        "linalg_summary": S(linalg='preconditioner = "None"\nverbosity = "Summary"\n'),
        # value conversions (toml++ node::value rules)
        "integers_as_numbers": S(initial="rho = 1\npressure = 101325\nvelocity = [1000, 0, -3]\n",
                                 time='mode = "Steady"\nscheme = "FORWARD_EULER"\ncfl = 1\nmax_iterations = 7\n'
                                 'residual_tolerance = 1\n'),
        "floats_as_integers": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 2.0\nmax_iterations = 500.0\n'
                                'residual_tolerance = 1e-8\n',
                                output='directory = "o"\nfield_interval = 0.0\nresidual_interval = 1e1\n'),
        "bool_as_integer": S(output='directory = "o"\nfield_interval = true\nresidual_interval = true\n'),
        "opt_number_wrong_type_default": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\n'
                                           'venkat_k = "big"\n', linalg='rel_tol = "x"\nmax_iter = 12.5\n'),
        "opt_string_wrong_type_default": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\n'
                                           'gradient = 3\n', linalg='solver = 1\npreconditioner = 2\nverbosity = false\n'),
        "hex_oct_bin_underscores": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 1_0.0_1\n'
                                     'max_iterations = 0x1_F4\nresidual_tolerance = 1e-1_0\n',
                                     output='directory = "o"\nfield_interval = 0o17\nresidual_interval = 0b101\n'),
        "strings_escapes_literal": S(output="directory = 'C:\\out\\dir'\nfield_interval = 1\nresidual_interval = 1\n"),
        "basic_string_escapes": S(output='directory = "out/\\u00e9t\\u00E9 \\"q\\" \\\\ tab\\t"\nfield_interval = 1\n'
                                  'residual_interval = 1\n'),
        "inline_tables_and_comments": (
            "# comment line\nflow = { eos = \"IDEAL_GAS\", flow_model = \"INVISCID_FLOW\", gamma = 1.4, "
            "gas_constant = 287.0 } # trailing\n"
            "initial = {rho = 1.0, pressure = 1e5, velocity = [ 1.0 , 2.0 ,\n 3.0, ]}\n"
            "[numerics] # section comment\nflux = \"HLLC\"\nreconstruction = \"FIRST_ORDER\"\n"
            "[time]\nmode = \"Steady\"\nscheme = \"BACKWARD_EULER\"\ncfl = +3.5\nmax_iterations = +10\n"
            "residual_tolerance = 5E-7\n"
            "[output]\n\"directory\" = \"quoted key\"\nfield_interval = 0\nresidual_interval = 1\n"),
        "dotted_keys": ("flow.eos = \"IDEAL_GAS\"\nflow.flow_model = \"INVISCID_FLOW\"\nflow.gamma = 1.3\n"
                        "flow.gas_constant = 300.0\n"
                        + "\n".join(f"[{k}]\n{v}" for k, v in BASE.items() if k != "flow")),
        "multiline_strings": S(output='directory = """\nmulti\\\n   line"""\nfield_interval = 1\nresidual_interval = 1\n'),
        "large_integer_exact": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 9007199254740992\n'
                                 'max_iterations = 9223372036854775807\nresidual_tolerance = 1e-3\n'),
        "date_values_are_not_numbers": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\n'
                                         'cfl = 1979-05-27T07:32:00Z\nmax_iterations = 3\nresidual_tolerance = 1e-3\n'),
        "date_value_as_string": S(output='directory = 1979-05-27\nfield_interval = 1\nresidual_interval = 1\n'),
        "time_value_as_default": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\n'
                                   'venkat_k = 07:32:00\n'),
        "inf_cfl": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = inf\nmax_iterations = 3\n'
                     'residual_tolerance = 1e-3\n'),
        # rejected files
        "unknown_root_key": S(_raw_x="[extra]\na = 1\n"),
        "unknown_root_keys_sorted": S(_raw_x="[zeta]\na = 1\n[alpha]\nb = 2\n"),
        "missing_flow": S(drop=("flow",)),
        "missing_initial": S(drop=("initial",)),
        "missing_numerics": S(drop=("numerics",)),
        "missing_time": S(drop=("time",)),
        "missing_output": S(drop=("output",)),
        "flow_not_table": S(flow=None, _raw_f='flow = 3\n'),
        "missing_eos": S(flow='flow_model = "INVISCID_FLOW"\ngamma = 1.4\ngas_constant = 287.0\n'),
        "eos_not_string": S(flow='eos = 1\nflow_model = "INVISCID_FLOW"\ngamma = 1.4\ngas_constant = 287.0\n'),
        "missing_flow_model": S(flow='eos = "IDEAL_GAS"\ngamma = 1.4\ngas_constant = 287.0\n'),
        "unsupported_eos": S(flow='eos = "REAL_GAS"\nflow_model = "INVISCID_FLOW"\n'),
        "missing_gamma": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngas_constant = 287.0\n'),
        "gamma_string": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = "1.4"\ngas_constant = 287.0\n'),
        "gamma_zero": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = 0\ngas_constant = 287.0\n'),
        "gas_constant_negative": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = 1.4\n'
                                   'gas_constant = -287.5\n'),
        "gamma_nan": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = nan\ngas_constant = 287.0\n'),
        "unsupported_flow_model": S(flow='eos = "IDEAL_GAS"\nflow_model = "RANS"\ngamma = 1.4\ngas_constant = 287.0\n'),
        "viscous_missing_prandtl": S(flow='eos = "IDEAL_GAS"\nflow_model = "VISCOUS_FLOW"\ngamma = 1.4\n'
                                     'gas_constant = 287.0\n'),
        "prandtl_negative": S(flow='eos = "IDEAL_GAS"\nflow_model = "VISCOUS_FLOW"\nprandtl = -1\ngamma = 1.4\n'
                              'gas_constant = 287.0\n'),
        "prandtl_in_inviscid": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\nprandtl = 0.7\ngamma = 1.4\n'
                                 'gas_constant = 287.0\n'),
        "flow_unknown_keys_sorted": S(flow='eos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = 1.4\n'
                                      'gas_constant = 287.0\nzz = 1\nbb = 2\n'),
        "initial_unknown_key": S(initial="rho = 1.0\npressure = 1e5\nvelocity = [1.0, 0.0, 0.0]\nmach = 2\n"),
        "rho_negative": S(initial="rho = -1.0\npressure = 1e5\nvelocity = [1.0, 0.0, 0.0]\n"),
        "pressure_zero": S(initial="rho = 1.0\npressure = 0\nvelocity = [1.0, 0.0, 0.0]\n"),
        "velocity_two": S(initial="rho = 1.0\npressure = 1e5\nvelocity = [1.0, 0.0]\n"),
        "velocity_string_element": S(initial='rho = 1.0\npressure = 1e5\nvelocity = [1.0, "a", 0.0]\n'),
        "velocity_not_array": S(initial="rho = 1.0\npressure = 1e5\nvelocity = 3.0\n"),
        "missing_velocity": S(initial="rho = 1.0\npressure = 1e5\n"),
        "numerics_unknown_key": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\nviscous = true\n'),
        "unsupported_flux": S(numerics='flux = "ROE"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\n'),
        "unsupported_reconstruction": S(numerics='flux = "HLLC"\nreconstruction = "WENO"\n'),
        "muscl_missing_limiter": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\n'),
        "muscl_bad_limiter": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "MINMOD_1D"\n'),
        "directional_bad_limiter": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL_DIRECTIONAL"\nlimiter = "VENKAT"\n'),
        "venkat_k_zero": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\nvenkat_k = 0.0\n'),
        "unsupported_gradient": S(numerics='flux = "HLLC"\nreconstruction = "MUSCL"\nlimiter = "VENKAT"\n'
                                  'gradient = "LEAST_SQUARES"\n'),
        "turbulence_not_table": S(flow=visc, _raw_t='turbulence = "SA"\n'),
        "turbulence_unknown_model": S(flow=visc, turbulence='model = "KOMEGA"\n'),
        "turbulence_missing_model": S(flow=visc, turbulence='nu_inf_ratio = 3\n'),
        "turbulence_unknown_key": S(flow=visc, turbulence='model = "SA"\nwall_functions = true\n'),
        "turbulence_inviscid": S(turbulence='model = "SA"\n'),
        "turbulence_ratio_zero": S(flow=visc, turbulence='model = "SA"\nnu_inf_ratio = 0\n'),
        "turbulence_tolerance_negative": S(flow=visc, turbulence='model = "SA"\ndistance_tolerance = -1e-8\n'),
        "time_unknown_key": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 25.0\nmax_iterations = 500\n'
                              'residual_tolerance = 1.0e-10\nstart_cfl = 1\n'),
        "unsupported_scheme": S(time='mode = "Steady"\nscheme = "RK4"\ncfl = 25.0\nmax_iterations = 500\n'
                                'residual_tolerance = 1.0e-10\n'),
        "cfl_negative": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = -1\nmax_iterations = 500\n'
                          'residual_tolerance = 1.0e-10\n'),
        "max_iterations_zero": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 1\nmax_iterations = 0\n'
                                 'residual_tolerance = 1.0e-10\n'),
        "max_iterations_fraction": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 1\nmax_iterations = 2.5\n'
                                     'residual_tolerance = 1.0e-10\n'),
        "residual_tolerance_zero": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 1\nmax_iterations = 5\n'
                                     'residual_tolerance = 0.0\n'),
        "missing_mode": S(time='scheme = "BACKWARD_EULER"\ncfl = 1\nmax_iterations = 5\nresidual_tolerance = 1e-3\n'),
        "unsteady_missing_time_step": S(time='mode = "Unsteady"\nscheme = "BACKWARD_EULER"\ncfl = 1\nmax_iterations = 5\n'
                                        'residual_tolerance = 1e-3\nmax_time_steps = 3\n'),
        "unsteady_missing_steps": S(time='mode = "Unsteady"\nscheme = "BACKWARD_EULER"\ncfl = 1\nmax_iterations = 5\n'
                                    'residual_tolerance = 1e-3\ntime_step = 1e-3\n'),
        "integer_beyond_double": S(initial="rho = 9007199254740993\npressure = 1e5\nvelocity = [1.0, 0.0, 0.0]\n"),
        "output_unknown_key": S(output='directory = "o"\nfield_interval = 1\nresidual_interval = 1\nformat = "vtu"\n'),
        "output_missing_directory": S(output='field_interval = 1\nresidual_interval = 1\n'),
        "output_field_negative": S(output='directory = "o"\nfield_interval = -1\nresidual_interval = 1\n'),
        "output_residual_zero": S(output='directory = "o"\nfield_interval = 1\nresidual_interval = 0\n'),
        "output_interval_string": S(output='directory = "o"\nfield_interval = "1"\nresidual_interval = 1\n'),
        "linalg_not_table": S(_raw_l='linalg = 1\n'),
        "linalg_unknown_key": S(linalg='solver = "BICGSTAB"\nrestart = 30\n'),
        "linalg_unsupported_solver": S(linalg='solver = "GMRES"\n'),
        "linalg_unsupported_precond": S(linalg='preconditioner = "ILU0"\n'),
        "linalg_unsupported_verbosity": S(linalg='verbosity = "Loud"\n'),
        "linalg_rel_tol_zero": S(linalg='rel_tol = 0.0\n'),
        "linalg_abs_tol_negative": S(linalg='abs_tol = -1e-9\n'),
        "linalg_max_iter_zero": S(linalg='max_iter = 0\n'),
    }
    syntax = {
        "syntax_missing_value": S(_raw_s="bad =\n"),
        "syntax_duplicate_key": S(initial="rho = 1.0\nrho = 2.0\npressure = 1e5\nvelocity = [1.0, 0.0, 0.0]\n"),
        "syntax_duplicate_table": S(_raw_s="[flow]\neos = \"IDEAL_GAS\"\n"),
        "syntax_unterminated_string": S(output='directory = "o\nfield_interval = 1\nresidual_interval = 1\n'),
        "syntax_bad_number": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 1.2.3\nmax_iterations = 5\n'
                               'residual_tolerance = 1e-3\n'),
        "syntax_leading_zero": S(time='mode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 012\nmax_iterations = 5\n'
                                 'residual_tolerance = 1e-3\n'),
        "syntax_unclosed_header": S(_raw_s="[extra\n"),
        "syntax_bad_date": S(_raw_s="when = 1979-05-27T25:61:00\n"),
    }
    return c, syntax


BC_HEAD = '[[boundary_condition]]\npatch_id = {pid}\nname = "{name}"\ncgns_type = "UserDefined"\nglobal_face_count = 10\n'


def bc_text(entries, names):
    out = []
    for pid, body in entries:
        out.append(BC_HEAD.format(pid=pid, name=names[pid] if 0 <= pid < len(names) else "x") + body)
    return "\n".join(out)


def bc_cases():
    names = ["inlet", "outlet", "wall", "top", "side", "far", "sub_in", "sub_out", "heat"]
    full = [
        (0, 'type = "SUPERSONIC_INLET"\np = 5590.34\nt = 107.14\nvelocity = [622.13, 0.0, 0.0]\n'),
        (1, 'type = "SUPERSONIC_OUTLET"\n'),
        (2, 'type = "NO_SLIP_WALL"\nt_wall = 300\nvelocity = [1.0, 2, 3.5]\n'),
        (3, 'type = "SLIP_WALL"\n'),
        (4, 'type = "SYMMETRY"\n'),
        (5, 'type = "FARFIELD"\nmach = 0.8\np_inf = 101325\nt_inf = 288.15\nalpha = 2.5\nbeta_deg = -1\n'),
        (6, 'type = "SUBSONIC_INLET"\nt = 290\nmach = 0.3\ndirection = [1.0, 1.0, 0.0]\n'),
        (7, 'type = "SUBSONIC_OUTLET"\np_back = 98000.0\n'),
        (8, 'type = "NO_SLIP_WALL_HEAT_FLUX"\ntmp_grad = -12.5\n'),
    ]
    ok = {
        "all_types": (full, names),
        "reverse_order": (list(reversed(full)), names),
        "velocity_inf_and_inf_keys": ([(0, 'type = "SUPERSONIC_INLET"\np_inf = 1e5\nt_inf = 300\n'
                                      'velocity_inf = [100, 0, 0]\n')], ["in"]),
        "farfield_direction": ([(0, 'type = "FARFIELD"\nmach = 2.0\np = 1e5\nt = 250\ndirection = [0, 1, 0]\n')], ["f"]),
        "farfield_velocity": ([(0, 'type = "FARFIELD"\np = 1e5\nt = 250\nvelocity = [10, 20, 30]\n')], ["f"]),
        "farfield_angles_default": ([(0, 'type = "FARFIELD"\nmach = 0.5\np = 1e5\nt = 250\n')], ["f"]),
        "farfield_alpha_deg_wins": ([(0, 'type = "FARFIELD"\nmach = 0.5\np = 1e5\nt = 250\nalpha = 1\nalpha_deg = 4\n'
                                      'beta = 2\n')], ["f"]),
        "no_slip_defaults": ([(0, 'type = "NO_SLIP_WALL"\n')], ["w"]),
        "no_slip_t_key": ([(0, 'type = "NO_SLIP_WALL"\nt = 310.0\n')], ["w"]),
        "adiabatic_alias": ([(0, 'type = "NO_SLIP_WALL_ADIABATIC"\nvelocity = [0, 0, 1]\n')], ["w"]),
        "heat_flux_grad_key": ([(0, 'type = "NO_SLIP_WALL_HEAT_FLUX"\nheat_flux_grad = 4\n')], ["w"]),
        "subsonic_inlet_default_p": ([(0, 'type = "SUBSONIC_INLET"\nt = 300\nvelocity = [50, 0, 0]\n')], ["s"]),
        "subsonic_inlet_angles": ([(0, 'type = "SUBSONIC_INLET"\np_inf = 9e4\nt_inf = 300\nmach = 0.2\nbeta = 3\n')], ["s"]),
        "subsonic_outlet_keys": ([(0, 'type = "SUBSONIC_OUTLET"\np_outlet = 1.5e5\n'), (1, 'type = "SUBSONIC_OUTLET"\np = 2e5\n'),
                                  (2, 'type = "SUBSONIC_OUTLET"\np_back = 1e5\np = 3\n')], ["a", "b", "c"]),
    }
    bad = {
        "no_entries": ("", ["a"]),
        "single_table_not_array": ('[boundary_condition]\npatch_id = 0\ntype = "SYMMETRY"\n', ["a"]),
        "root_unknown_key": ('version = 2\n' + bc_text([(0, 'type = "SYMMETRY"\n')], ["a"]), ["a"]),
        "entry_not_table": ('boundary_condition = [1, 2]\n', ["a"]),
        "patch_id_missing": ('[[boundary_condition]]\ntype = "SYMMETRY"\n', ["a"]),
        "patch_id_string": ('[[boundary_condition]]\npatch_id = "0"\ntype = "SYMMETRY"\n', ["a"]),
        "patch_id_out_of_range": (bc_text([(3, 'type = "SYMMETRY"\n')], ["a"]), ["a"]),
        "patch_id_negative": (bc_text([(-1, 'type = "SYMMETRY"\n')], ["a"]), ["a"]),
        "duplicate_patch": (bc_text([(0, 'type = "SYMMETRY"\n'), (0, 'type = "SLIP_WALL"\n')], ["a"]), ["a"]),
        "unassigned_patch": (bc_text([(0, 'type = "SYMMETRY"\n')], ["a", "b_patch"]), ["a", "b_patch"]),
        "unknown_type": (bc_text([(0, 'type = "PERIODIC"\n')], ["a"]), ["a"]),
        "missing_type": ('[[boundary_condition]]\npatch_id = 0\n', ["a"]),
        "symmetry_extra_key": (bc_text([(0, 'type = "SYMMETRY"\np = 1e5\n')], ["a"]), ["a"]),
        "velocity_and_mach": (bc_text([(0, 'type = "SUPERSONIC_INLET"\np = 1\nt = 1\nvelocity = [1, 0, 0]\nmach = 2\n')],
                                      ["a"]), ["a"]),
        "direction_and_angles": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\nmach = 2\ndirection = [1, 0, 0]\n'
                                            'alpha = 3\n')], ["a"]), ["a"]),
        "direction_without_mach": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\ndirection = [1, 0, 0]\n')], ["a"]), ["a"]),
        "angles_without_mach": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\nbeta_deg = 3\n')], ["a"]), ["a"]),
        "no_velocity_definition": (bc_text([(0, 'type = "SUPERSONIC_INLET"\np = 1\nt = 1\n')], ["a"]), ["a"]),
        "mach_inf_not_whitelisted": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\nmach_inf = 2\n')], ["a"]), ["a"]),
        "missing_pressure": (bc_text([(0, 'type = "SUPERSONIC_INLET"\nt = 1\nvelocity = [1, 0, 0]\n')], ["a"]), ["a"]),
        "missing_temperature": (bc_text([(0, 'type = "SUPERSONIC_INLET"\np = 1\nvelocity = [1, 0, 0]\n')], ["a"]), ["a"]),
        "pressure_negative": (bc_text([(0, 'type = "SUPERSONIC_INLET"\np = -1\nt = 1\nvelocity = [1, 0, 0]\n')], ["a"]),
                              ["a"]),
        "temperature_zero": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt_inf = 0\nmach = 1\n')], ["a"]), ["a"]),
        "mach_zero": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\nmach = 0\n')], ["a"]), ["a"]),
        "direction_zero": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\nmach = 2\ndirection = [0, 0, 1e-8]\n')], ["a"]),
                           ["a"]),
        "direction_two": (bc_text([(0, 'type = "FARFIELD"\np = 1\nt = 1\nmach = 2\ndirection = [1, 0]\n')], ["a"]), ["a"]),
        "subsonic_outlet_missing": (bc_text([(0, 'type = "SUBSONIC_OUTLET"\n')], ["a"]), ["a"]),
        "subsonic_outlet_negative": (bc_text([(0, 'type = "SUBSONIC_OUTLET"\np_back = -5\n')], ["a"]), ["a"]),
        "subsonic_outlet_extra": (bc_text([(0, 'type = "SUBSONIC_OUTLET"\np_back = 5\nt = 3\n')], ["a"]), ["a"]),
        "no_slip_t_negative": (bc_text([(0, 'type = "NO_SLIP_WALL"\nt_wall = -3\n')], ["a"]), ["a"]),
        "no_slip_extra_key": (bc_text([(0, 'type = "NO_SLIP_WALL"\ntmp_grad = 3\n')], ["a"]), ["a"]),
        "heat_flux_extra_key": (bc_text([(0, 'type = "NO_SLIP_WALL_HEAT_FLUX"\nt_wall = 300\n')], ["a"]), ["a"]),
        "subsonic_inlet_velocity_and_mach": (bc_text([(0, 'type = "SUBSONIC_INLET"\nt = 1\nmach = 0.2\n'
                                                        'velocity = [1, 0, 0]\n')], ["a"]), ["a"]),
    }
    syntax = {
        "syntax_bc_bad_array": (bc_text([(0, 'type = "SYMMETRY"\n')], ["a"]) + "x = [1, 2\n", ["a"]),
    }
    return ok, bad, syntax, names


def parse_output(stdout):
    """The harness's "key value" lines: integers, doubles (17 digits) and the output directory."""
    out = {}
    for ln in stdout.splitlines():
        k, _, v = ln.partition(" ")
        if k == "output_dir":
            out[k] = v
        else:
            try:
                out[k] = int(v)
            except ValueError:
                out[k] = float(v)
    return out


def run(kind, text, names, tmp):
    path = "case.toml"
    with open(os.path.join(tmp, path), "w", encoding="utf-8") as f:
        f.write(text)
    args = [HARNESS, kind, path] + (list(names) if kind == "bc" else [])
    p = subprocess.run(args, cwd=tmp, capture_output=True, text=True, timeout=60)
    warnings = [ln for ln in p.stderr.splitlines() if ln.startswith("WARNING: ")]
    if p.returncode == 0:
        return {"expect": parse_output(p.stdout), "warnings": warnings}
    marker = "CGNS Reader Fatal Error: "
    msgs = [ln.split(marker, 1)[1] for ln in p.stderr.splitlines() if marker in ln]
    if not msgs:
        raise RuntimeError(f"reference parser failed without a message: {p.stderr}")
    return {"error": msgs[0]}


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_config missing: run oracle/build_ref.sh")
    cases = []
    with tempfile.TemporaryDirectory() as tmp:
        sc, ss = solver_cases()
        for name, text in sc.items():
            cases.append(dict(name=name, kind="solver", text=text, **run("solver", text, [], tmp)))
        for name, text in ss.items():
            r = run("solver", text, [], tmp)
            assert "error" in r, name
            cases.append(dict(name=name, kind="solver", text=text, syntax=True, **r))
        ok, bad, syn, _ = bc_cases()
        for name, (entries, names) in ok.items():
            text = bc_text(entries, names)
            r = run("bc", text, names, tmp)
            assert "expect" in r, (name, r)
            cases.append(dict(name=name, kind="bc", text=text, patch_names=names, **r))
        for name, (text, names) in bad.items():
            r = run("bc", text, names, tmp)
            assert "error" in r, (name, r)
            cases.append(dict(name=name, kind="bc", text=text, patch_names=names, **r))
        for name, (text, names) in syn.items():
            r = run("bc", text, names, tmp)
            assert "error" in r, name
            cases.append(dict(name=name, kind="bc", text=text, patch_names=names, syntax=True, **r))
    # the reference's example files: parsed values only
    examples = {}
    for f in sorted(glob.glob(os.path.join(REF, "examples", "*", "*.toml"))):
        rel = os.path.relpath(f, REF)
        if os.path.basename(f).startswith("bc"):
            names = patch_names_of(f)
            p = subprocess.run([HARNESS, "bc", f] + names, capture_output=True, text=True, timeout=60)
            kind = "bc"
        else:
            names = []
            p = subprocess.run([HARNESS, "solver", f], capture_output=True, text=True, timeout=60)
            kind = "solver"
        if p.returncode != 0:
            raise RuntimeError(f"{rel}: {p.stderr}")
        examples[rel] = dict(kind=kind, patch_names=names, expect=parse_output(p.stdout))
    with open(os.path.join(HERE, "config_cases.json"), "w", encoding="utf-8") as f:
        json.dump({"cases": cases, "ref_examples": examples}, f, indent=1, sort_keys=True)
    print(f"wrote {len(cases)} cases and {len(examples)} reference examples")


def patch_names_of(path):
    """Patch names of a bc file in patch_id order (its 'name' entries)."""
    names = {}
    pid, name = None, None
    with open(path, encoding="utf-8") as f:
        for ln in f:
            s = ln.split("#", 1)[0].strip()
            if s.startswith("[[boundary_condition]]"):
                if pid is not None:
                    names[pid] = name
                pid, name = None, None
            elif s.startswith("patch_id"):
                pid = int(s.split("=", 1)[1])
            elif s.startswith("name"):
                name = s.split("=", 1)[1].strip().strip('"')
    if pid is not None:
        names[pid] = name
    return [names[k] or f"patch{k}" for k in sorted(names)]


if __name__ == "__main__":
    main()
