"""Output files after GPU steps against the same files formatted from the oracle's
state with the reference's formats.

The writers (class_system::print_output, src/system_print_output.cpp:14-80)
print ID order: micro-<step>.dat as %d then REAL_FMT "%.8lf" columns
(2D: id x y vx vy m d p type cxx cxy cyy, with a space before the newline;
3D: id x y z vx vy vz m d p type cxx cxy cyy cxz cyz czz;
src/system_print_particles.cpp:57-91), walls-<step>.dat as "i fx fy [fz]"
(src/system_print_walls.cpp:32-44), colloids-<step>.dat (2D: i x y vx vy wz fx fy
tz; 3D: i x y z vx vy vz wx wy wz fx fy fz tx ty tz;
src/system_print_colloids.cpp:81-111) and macro.dat lines "%d %f" of the step and
the kinetic energy (src/system_print_macro_vars.cu:37-47).

Checked: the file set, the line count, the exact layout of every line (integer
columns, 8 decimals per real column, single spaces, the 2D trailing space),
integer columns equal, and real columns within 1.5e-8 absolute of the oracle's
(the printed resolution: FP64 state agrees to ~1e-12 relative, so a printed
value moves by at most one unit of its 8th decimal) - 1.5e-6 for "%f".
"""
import os
import re

import numpy as np
import pytest

import helpers
import oracle
import sph_system

pytestmark = pytest.mark.gpu

R8 = r"-?\d+\.\d{8}"


def _micro_ref(o, dim):
    cols = (["x", "y", "vx", "vy", "mass", "dens", "press", "type", "cxx", "cxy", "cyy"]
            if dim == 2 else
            ["x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "type", "cxx", "cxy",
             "cyy", "cxz", "cyz", "czz"])
    data = [o.get(c) for c in cols]
    lines = []
    for i in range(o.params.N):
        toks = [str(i)] + [str(int(v[i])) if c == "type" else "%.8f" % v[i]
                           for c, v in zip(cols, data)]
        lines.append(" ".join(toks) + (" " if dim == 2 else ""))
    pattern = (r"\d+" + "".join(r" \d+" if c == "type" else " " + R8 for c in cols)
               + (" " if dim == 2 else ""))
    return lines, pattern


def _walls_ref(o, dim):
    f = [o.get(n) for n in ("fx_wall", "fy_wall", "fz_wall")[:dim]]
    lines = [" ".join([str(w)] + ["%.8f" % a[w] for a in f]) for w in range(o.params.Nwalls)]
    return lines, r"\d+" + (" " + R8) * dim


def _colloids_ref(o, dim):
    cols = (["coll_x", "coll_y", "coll_vx", "coll_vy", "coll_omegaz", "fx_colloid", "fy_colloid",
             "tz_colloid"] if dim == 2 else
            ["coll_x", "coll_y", "coll_z", "coll_vx", "coll_vy", "coll_vz", "coll_omegax",
             "coll_omegay", "coll_omegaz", "fx_colloid", "fy_colloid", "fz_colloid",
             "tx_colloid", "ty_colloid", "tz_colloid"])
    data = [o.get(c) for c in cols]
    lines = [" ".join([str(i)] + ["%.8f" % v[i] for v in data])
             for i in range(o.params.N_colloids)]
    return lines, r"\d+" + (" " + R8) * len(cols)


def _compare(path, ref, atol=1.5e-8):
    lines, pattern = ref
    with open(path) as f:
        text = f.read()
    assert text.endswith("\n"), path
    got = text[:-1].split("\n")
    assert len(got) == len(lines), (path, len(got), len(lines))
    rx = re.compile(pattern + r"\Z")
    for k, (g, r) in enumerate(zip(got, lines)):
        assert rx.match(g), f"{path} line {k}: layout {g!r}"
        gt, rt = g.split(), r.split()
        for a, b in zip(gt, rt):
            if "." in b:
                assert abs(float(a) - float(b)) <= atol, f"{path} line {k}: {a} vs {b}"
            else:
                assert a == b, f"{path} line {k}: {a} vs {b}"


CASES = {
    "colloids2d": lambda: helpers.cfg_colloids_2d(n=48),
    "colloids3d": lambda: helpers.cfg_colloids_3d(n=20),
}


@pytest.mark.parametrize("case", list(CASES))
def test_print_output_after_steps(case, tmp_path):
    cfg = CASES[case]()
    cfg.update(freq_micro=5, freq_walls=5, freq_colloids=10, freq_macro=2)
    s = helpers.make_system(cfg, str(tmp_path))
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    s.prepare()
    o.prepare()
    s.step(10)
    o.step(10)
    out = str(tmp_path / "out")
    os.makedirs(out)
    s.print_output(10, out)
    dim = s.params.dim
    assert sorted(os.listdir(out)) == ["colloids-10.dat", "macro.dat", "micro-10.dat",
                                       "walls-10.dat"]
    _compare(os.path.join(out, "micro-10.dat"), _micro_ref(o, dim))
    _compare(os.path.join(out, "walls-10.dat"), _walls_ref(o, dim))
    _compare(os.path.join(out, "colloids-10.dat"), _colloids_ref(o, dim))
    ke = o.kinetic_energy()
    _compare(os.path.join(out, "macro.dat"), (["10 %f" % ke], r"10 -?\d+\.\d{6}"), atol=1.5e-6)
    # a step that is a multiple of freq_macro only appends to macro.dat
    s.step(2)
    o.step(2)
    s.print_output(12, out)
    assert sorted(os.listdir(out)) == ["colloids-10.dat", "macro.dat", "micro-10.dat",
                                       "walls-10.dat"]
    _compare(os.path.join(out, "macro.dat"),
             (["10 %f" % ke, "12 %f" % o.kinetic_energy()], r"1[02] -?\d+\.\d{6}"),
             atol=1.5e-6)


def test_run_writes_the_reference_file_set(tmp_path):
    """sph_run: prepare, output at step 0, Nsteps steps with output at the
    freq_* multiples (src/main.cu:30-561); the last micro file and every
    macro.dat line against the oracle driven through the same loop."""
    cfg = helpers.cfg_couette(n=32, steps=20)
    cfg.update(freq_micro=10, freq_walls=10, freq_colloids=10, freq_macro=5)
    s = helpers.make_system(cfg, str(tmp_path))
    o = oracle.Oracle.from_system(s)
    out = str(tmp_path / "run")
    os.makedirs(out)
    s.run(out)
    assert sorted(os.listdir(out)) == sorted(
        ["macro.dat"] + [f"micro-{k}.dat" for k in (0, 10, 20)]
        + [f"walls-{k}.dat" for k in (0, 10, 20)])
    o.prepare()
    ke = [o.kinetic_energy()]
    walls0 = _walls_ref(o, 2)
    for _ in range(4):
        o.step(5)
        ke.append(o.kinetic_energy())
    _compare(os.path.join(out, "micro-20.dat"), _micro_ref(o, 2))
    _compare(os.path.join(out, "walls-20.dat"), _walls_ref(o, 2))
    _compare(os.path.join(out, "walls-0.dat"), walls0)
    _compare(os.path.join(out, "macro.dat"),
             (["%d %f" % (5 * k, e) for k, e in enumerate(ke)], r"\d+ -?\d+\.\d{6}"),
             atol=1.5e-6)
