"""The Fortran drop-in surface: modules cart_mpi / initialization / lbm over
the C ABI, the driver with main.f90's call sequence, and the output file.

CPU: the reference's own main.f90 compiles and links unchanged against the
modules (when the reference sources are present); the random_number fixture
equals the values the Fortran runtime is known to produce.
GPU: the driver's Monitor lines and output_<count>.plt equal the oracle's run
of the same deck (bit for bit for the fields, and for Umax to the printed
digits), and the file header matches the reference's Tecplot layout.
"""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

import imexlb as I
from cases import TINY, get_deck, oracle_setup

FORTRAN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "imexlb-1.0_amd", "fortran")
REF_MAIN = "/root/reference/main.f90"
FLANG = "/opt/rocm/bin/amdflang"


def test_r_vec_fixture_is_the_runtime_sequence(r_vec):
    # first five values of an unseeded random_number(r) in this runtime
    first = [2.24780151922288063e-05, 6.01352605402107554e-01, 9.67955701695466186e-01,
             5.14975824298344875e-01, 2.62906165554665394e-01]
    assert np.array_equal(r_vec[:5], first)
    assert np.all((r_vec >= 0) & (r_vec < 1)) and len(set(r_vec.tolist())) == 100


@pytest.mark.skipif(not (os.path.exists(REF_MAIN) and os.path.exists(FLANG)), reason="reference sources absent")
@pytest.mark.parametrize("variant", ["2d", "3d"])
def test_reference_main_links_against_modules(variant, tmp_path):
    build = os.path.join(FORTRAN, "build", variant)
    if not os.path.exists(os.path.join(build, "lbm.mod")):
        pytest.skip("Fortran modules not built")
    obj = tmp_path / "main.o"
    subprocess.run([FLANG, "-O2", "-I", build, "-c", REF_MAIN, "-o", str(obj)], check=True)
    objs = [os.path.join(build, m) for m in ("mpi_mod.o", "imexlb_capi.o", "cart_mpi.o", "initialization.o",
                                            "lbm.o")]
    lib = os.path.dirname(FORTRAN)
    subprocess.run([FLANG, str(obj), *objs, f"-L{lib}", "-limexlb", f"-Wl,-rpath,{lib}",
                    "/opt/conda/lib/libmpifort.so", "/opt/conda/lib/libmpi.so", "-Wl,-rpath,/opt/conda/lib",
                    "-o", str(tmp_path / "run")], check=True)
    assert (tmp_path / "run").exists()


def read_plt(path, dim):
    """Parse the Tecplot binary the reference writes (lbm.f90:258-390)."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"#!TDV112"
    pos = 8
    numvars = 5 if dim == 2 else 7  # coordinates + p + velocity components
    title = "Cylinder" if dim == 2 else "Sphere"
    names = ["x", "y", "p", "ux", "uy"] if dim == 2 else ["x", "y", "z", "p", "ux", "uy", "uz"]
    ints = [1, 0] + [ord(c) for c in title] + [0, numvars]
    for nm in names:
        ints += [ord(c) for c in nm] + [0]
    got = struct.unpack_from(f"<{len(ints)}i", raw, pos)
    assert list(got) == ints
    pos += 4 * len(ints)
    assert struct.unpack_from("<f", raw, pos)[0] == 299.0
    pos += 4
    zone = struct.unpack_from("<11i", raw, pos)
    pos += 44
    solution_time = struct.unpack_from("<d", raw, pos)[0]
    pos += 8
    dims = struct.unpack_from("<9i", raw, pos)
    pos += 36
    assert struct.unpack_from("<2f", raw, pos) == (357.0, 299.0)
    pos += 8
    fmt = struct.unpack_from(f"<{numvars + 3}i", raw, pos)
    assert list(fmt) == [2] * numvars + [0, 0, -1]
    pos += 4 * (numvars + 3)
    limits = np.frombuffer(raw, "<f8", 2 * numvars, pos)
    pos += 16 * numvars
    nodes = dims[5] * dims[6] * dims[7]
    data = np.frombuffer(raw, "<f8", numvars * nodes, pos).reshape(numvars, nodes)
    assert pos + 8 * numvars * nodes == len(raw)
    zone_name = "".join(chr(c) for c in zone[:8])
    return {"zone": zone_name, "zone_tail": zone[8:], "time": solution_time, "dims": dims, "limits": limits,
            "data": dict(zip(names, data))}


def _run_driver(variant, deck_text, workdir, env_extra=None):
    exe = os.path.join(FORTRAN, "build", f"imexlb{variant}")
    name = "input.in" if variant == "2d" else "input_3d.in"
    with open(os.path.join(workdir, name), "w") as fh:
        fh.write(deck_text)
    env = dict(os.environ)
    env.update(env_extra or {})
    print(f"[fortran] {exe} in {workdir} env={env_extra}", flush=True)
    out = subprocess.run([exe], cwd=workdir, env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.gpu
def test_driver_prints_forces(r_vec, tmp_path):
    """IMEXLB_FORCES=1 adds drag and lift to the Monitor lines; they equal the
    oracle's momentum exchange after the same number of steps."""
    deck = get_deck("tiny2d")
    stdout = _run_driver("2d", TINY["tiny2d"][1], str(tmp_path), {"IMEXLB_FORCES": "1"})
    d = I.derive(deck)
    drag = [float(v) for v in re.findall(r"Drag=\s*(\S+)", stdout)]
    lift = [float(v) for v in re.findall(r"Lift=\s*(\S+)", stdout)]
    o, _ = oracle_setup(deck, r_vec)
    expect = []
    for k in range(1, len(drag) + 1):
        o.run((k - 1) * d.interv, k * d.interv - 1, 0)
        expect.append(o.forces())
    expect = np.array(expect)
    assert len(drag) == (d.max_step // d.interv) and len(drag) > 3
    scale = np.abs(expect).max()
    assert np.all(np.abs(np.array(drag) - expect[:, 0]) <= 1e-12 * scale)
    assert np.all(np.abs(np.array(lift) - expect[:, 1]) <= 1e-12 * scale)
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case,variant,env", [("tiny2d", "2d", {}), ("tiny2d", "2d", {"IMEXLB_FUSED": "0"}),
                                              ("tiny3d", "3d", {})])
def test_driver_matches_oracle(case, variant, env, r_vec, tmp_path):
    deck = get_deck(case)
    stdout = _run_driver(variant, TINY[case][1], str(tmp_path), env)
    d = I.derive(deck)
    o, _ = oracle_setup(deck, r_vec)  # the driver draws r_vec with the same runtime
    um = o.run(0, d.max_step, d.interv)
    printed = [float(v) for v in re.findall(r"Umax=\s*(\S+)", stdout)]
    assert len(printed) == len(um) > 0
    assert np.allclose(printed, um / (d.charlength * d.t_intv), rtol=1e-14, atol=0)
    count = len(um)
    plt = read_plt(os.path.join(str(tmp_path), f"output_{count:03d}.plt"), deck.dim)
    assert plt["zone"] == f"ZONE {count:03d}" and plt["zone_tail"] == (0, -1, 0)
    assert plt["time"] == (d.max_step + 1) / d.t_intv  # iter after the loop (main.f90:48-73)
    nx, ny, nz = deck.n
    assert plt["dims"][5:8] == (nx + 1, ny + 1, (nz + 1) if deck.dim == 3 else 1)
    interior = o.interior(o.p, 1)[..., 0].ravel()
    assert np.array_equal(plt["data"]["p"], interior * d.charlength / d.t_intv ** 2)
    u = o.interior(o.u, deck.dim)
    for k, nm in enumerate(["ux", "uy", "uz"][:deck.dim]):
        assert np.array_equal(plt["data"][nm], u[..., k].ravel() / (d.charlength * d.t_intv))
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,rtol", [("f32", 1e-4), ("f16", 3e-2), ("bf16sr", 1e-1)])
def test_driver_storage_forms(prec, rtol, r_vec, tmp_path):
    """IMEXLB_PRECISION selects a storage form through the drop-in modules:
    the Monitor series of the 2D deck stays within rtol of the oracle's FP64
    series."""
    deck = get_deck("tiny2d")
    stdout = _run_driver("2d", TINY["tiny2d"][1], str(tmp_path), {"IMEXLB_PRECISION": prec})
    d = I.derive(deck)
    o, _ = oracle_setup(deck, r_vec)
    um = o.run(0, d.max_step, d.interv) / (d.charlength * d.t_intv)
    printed = np.array([float(v) for v in re.findall(r"Umax=\s*(\S+)", stdout)])
    assert len(printed) == len(um) > 3
    assert np.allclose(printed[1:], um[1:], rtol=rtol, atol=0), np.abs(printed / um - 1).max()
    o.close()


@pytest.mark.gpu
def test_driver_aborts_on_divergence(tmp_path):
    """A deck far past BGK's stability limit (inlet velocity about 70
    lattice units at ome 1.99) turns non-finite within ten steps: the first
    Monitor after that (iteration 17) reports IMEXLB_E_NAN and the drop-in's
    Monitor prints the error and calls MPI_ABORT, instead of printing a
    maximum that left the NaN out."""
    text = "48 24\n0.25 0.25 1000000\n1.99 1 1\n0.002 0.0001\n"
    d = I.derive(I.parse_deck(text, 2))
    assert d.interv == 17 and d.uu > 50
    exe = os.path.join(FORTRAN, "build", "imexlb2d")
    with open(os.path.join(str(tmp_path), "input.in"), "w") as fh:
        fh.write(text)
    print(f"[fortran] divergence deck in {tmp_path}", flush=True)
    out = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    text_out = out.stdout + out.stderr
    assert out.returncode != 0, text_out
    assert "Monitor failed" in text_out and "non-finite" in text_out, text_out
    assert len(re.findall(r"Umax=", text_out)) == 1  # only iteration 0 was printed
    assert not os.path.exists(os.path.join(str(tmp_path), "output_001.plt"))
