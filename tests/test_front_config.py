"""The configuration readers of the host front end against the reference's own
parsers.

tests/golden/config_cases.json holds configuration files written for this
test and what the reference's parse_solver_config / parse_boundary_config
(src/solver/compressible/config.cpp:28-288, src/bc/config.cpp:141-289, with
the vendored toml++ 3.4.0; run by tests/golden/make_config_golden.py) made of
each: the parsed structure, the log_stat warnings, or the message of the
mpi::fatal that rejected the file.  cfdg_config_read_solver / _boundary must
give the same structure bit for bit (doubles compared as bits), the same
warnings and the same message.  For a TOML syntax error only the message's
prefix "config: cannot parse '<path>': " is the reference's; the description
after it is the reader's own.  Where the reference tree is present, its own
example files are parsed too and compared with the values the reference's
parser gave for them.  Runs on the CPU (host code of libcfdg.so only).
"""
import json
import os
import struct

import pytest

import cfdg
import cfdg.front as fr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "config_cases.json"), encoding="utf-8"))
REF = "/root/reference"

SOLVER_KEYS = [
    "eos", "gamma", "gas_constant", "prandtl", "flow_model", "init_rho", "init_p", "gradient", "flux",
    "reconstruction", "limiter", "venkat_k", "turbulence", "nu_inf_ratio", "max_distance_sweeps",
    "distance_tolerance", "scheme", "cfl", "max_iterations", "residual_tolerance", "time_mode", "dt",
    "max_time_steps", "bdf_order", "solver", "preconditioner", "rel_tol", "abs_tol", "max_iter", "verbosity",
    "output_dir", "field_interval", "residual_interval",
]
BC_KEYS = ["type", "inflow_mode", "p", "t", "tmp_grad", "mach", "alpha_deg", "beta_deg"]


def same(a, b):
    """Equal values; doubles compared bit for bit (NaN included)."""
    if isinstance(a, float) or isinstance(b, float):
        return struct.pack("<d", float(a)) == struct.pack("<d", float(b))
    return a == b


def check_solver(cfg, expect, what):
    d = cfg.as_dict()
    for k in SOLVER_KEYS:
        assert same(d[k], expect[k]), (what, k, d[k], expect[k])
    for j in range(3):
        assert same(d["init_velocity"][j], expect[f"init_velocity{j}"]), (what, "init_velocity", j)


def check_bcs(bcs, expect, what):
    for p, b in enumerate(bcs):
        for k in BC_KEYS:
            assert same(getattr(b, k), expect[f"bc{p}.{k}"]), (what, p, k, getattr(b, k), expect[f"bc{p}.{k}"])
        for j in range(3):
            assert same(b.velocity[j], expect[f"bc{p}.velocity{j}"]), (what, p, "velocity", j)
            assert same(b.direction[j], expect[f"bc{p}.direction{j}"]), (what, p, "direction", j)


CASES = GOLD["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] + "-" + c["kind"] for c in CASES])
def test_case_matches_reference_parser(case):
    if case["kind"] == "solver":
        run = lambda: fr.read_solver_config("case.toml", text=case["text"])  # noqa: E731
    else:
        run = lambda: fr.read_boundary_config("case.toml", patch_names=case["patch_names"], text=case["text"])  # noqa: E731
    if "error" in case:
        with pytest.raises(fr.ConfigError) as ei:
            run()
        assert ei.value.code == cfdg.ERR_CONFIG
        msg = str(ei.value).split(": ", 1)[1]  # drop the "cfdg error -5" prefix
        if case.get("syntax"):
            prefix = "config: cannot parse 'case.toml': "
            assert case["error"].startswith(prefix) and msg.startswith(prefix), (msg, case["error"])
        else:
            assert msg == case["error"]
        return
    if case["kind"] == "solver":
        check_solver(run(), case["expect"], case["name"])
    else:
        bcs, warnings = run()
        check_bcs(bcs, case["expect"], case["name"])
        assert warnings == case["warnings"]


def test_every_reference_message_is_covered():
    """Each distinct failure the reference parsers can report appears in the
    fixture at least once (the message templates of config.cpp / bc/config.cpp
    / toml_utilities.cpp, with their variable parts removed)."""
    errors = " ".join(c["error"] for c in CASES if "error" in c)
    templates = [
        "unknown key", "missing required section", "missing required key", "must be a number", "must be an integer",
        "must be a string", "must be an array of 3 numbers", "must be positive", "unsupported EOS model",
        "unsupported Flow Model", "unsupported flux", "invalid limiter", "for MUSCL_DIRECTIONAL",
        "unsupported reconstruction", "unsupported gradient", "[turbulence] must be a table",
        "unsupported turbulence model", "turbulence requires", "unsupported scheme", "'max_iterations' must be >= 1",
        "intervals must be", "[linalg] must be a table", "unsupported solver type", "unsupported preconditioner type",
        "unsupported verbosity type", "cannot parse", "no [[boundary_condition]] entries", "must be a TOML table",
        "outside mesh range", "duplicate condition", "unknown BC type", "conflicting parameters: cannot specify direct",
        "conflicting parameters: cannot specify both", "'direction' vector requires", "'alpha/beta' angles require",
        "missing velocity definition", "missing required pressure", "missing required temperature",
        "'direction' vector cannot be zero", "missing required backpressure", "no condition assigned",
    ]
    for t in templates:
        assert t in errors, t


def test_file_and_text_readers_agree(tmp_path):
    case = next(c for c in CASES if c["name"] == "linalg_full")
    p = tmp_path / "case.toml"
    p.write_text(case["text"], encoding="utf-8")
    a = fr.read_solver_config(str(p)).as_dict()
    b = fr.read_solver_config(str(p), text=case["text"]).as_dict()
    assert a == b
    with pytest.raises(fr.ConfigError, match="config: cannot open file '.*missing.toml' on Rank 0"):
        fr.read_solver_config(str(tmp_path / "missing.toml"))


def test_boundary_file_against_a_mesh(tmp_path):
    m = cfdg.Mesh.generate(kind=0, ni=4, nj=3, nk=2, x0=0, x1=1, y0=0, y1=1, z0=0, z1=1)
    names = m.patch_names()
    text = "\n".join(f'[[boundary_condition]]\npatch_id = {p}\ntype = "SYMMETRY"\n' for p in range(len(names)))
    p = tmp_path / "bc.toml"
    p.write_text(text, encoding="utf-8")
    bcs, w = fr.read_boundary_config(str(p), mesh=m)
    assert [b.type for b in bcs] == [cfdg.BC["SYMMETRY"]] * len(names) and w == []
    # a missing patch is reported with the mesh's name for it
    p.write_text(text.rsplit("[[boundary_condition]]", 1)[0], encoding="utf-8")
    with pytest.raises(fr.ConfigError, match=f"no condition assigned to patch {len(names) - 1} \\('{names[-1]}'\\)"):
        fr.read_boundary_config(str(p), mesh=m)


def test_numerics_of_a_config():
    case = next(c for c in CASES if c["name"] == "linalg_full")
    num, lp = fr.to_numerics(fr.read_solver_config("case.toml", text=case["text"]))
    assert (num.reconstruction, num.limiter, num.gradient) == (cfdg.MUSCL, cfdg.LIMITER["VENKAT"],
                                                              cfdg.GRADIENT["GREEN_GAUSS_FACE"])
    assert (lp.preconditioner, lp.rel_tol, lp.abs_tol, lp.max_iter) == (cfdg.PRECOND_SGS, 1e-2, 1e-14, 250)
    assert num.prandtl_turb == 0.9 and num.cfl == 25.0
    for name, expect in (("van_albada_lsq_face", (cfdg.LIMITER["VAN_ALBADA"], cfdg.GRADIENT["LEAST_SQUARES_NODE"])),
                         ("viscous_barth_gg_cell", (cfdg.LIMITER["BARTH"], cfdg.GRADIENT["GREEN_GAUSS_CELL"]))):
        c = next(c for c in CASES if c["name"] == name)
        num, _ = fr.to_numerics(fr.read_solver_config("case.toml", text=c["text"]))
        assert (num.limiter, num.gradient) == expect
    c = next(c for c in CASES if c["name"] == "directional_minmod")
    with pytest.raises(fr.ConfigError, match="MUSCL_DIRECTIONAL reconstruction is not available"):
        fr.to_numerics(fr.read_solver_config("case.toml", text=c["text"]))


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="reference tree not present")
@pytest.mark.parametrize("rel", sorted(GOLD["ref_examples"]))
def test_reference_example_files(rel):
    ex = GOLD["ref_examples"][rel]
    path = os.path.join(REF, rel)
    if ex["kind"] == "solver":
        check_solver(fr.read_solver_config(path), ex["expect"], rel)
    else:
        bcs, _ = fr.read_boundary_config(path, patch_names=ex["patch_names"])
        check_bcs(bcs, ex["expect"], rel)
