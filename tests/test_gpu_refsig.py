"""The reference-signature boundary (include/ibtfo_ref.h) driven the way the
reference's test drivers drive it: tests/drivers/ref_sig_driver.cpp keeps the
Test/Sod/Sod.cpp call sequence (Psy_coord, Level_Set, Scheme_Index, Initialize,
Boundary, ComputeDt, then TimeAdvance / Boundary / Global_Boundary / ComputeDt /
CaculateMass per step) with the OpenACC data directives replaced by the four
ibtfo_ref_* calls, and reads its mask from a PNG named in the .inp (the PNGGeo
case). Its results must equal the oracle's: cell types, state and dt bit for
bit, mass within the oracle's sequential-summation bound."""
import os
import subprocess

import numpy as np
import pytest

from parity_util import nan_equal_bits

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "drivers", "ref_sig_driver")

INP = """N_x = {nx}
N_y = {ny}
num_ghost_cell = 2
lo_x = 0.0
lo_y = 0.0
hi_x = 2.0
hi_y = 1.0
Psy_time = {t_end!r}
max_iter = {steps}
clip_last_step = {clip}
call_advance = {adv}
CFL_number = 0.5
gamma = 1.4
Scheme = MUSCL
Riemann = {riemann}
Limiter = VanLeer
Boundary = {boundary}
Time = SingleStep
Edge = inflow
quirks = 0
inflow_rho = 1.4
inflow_u = 3.0
inflow_v = 0.0
inflow_p = 1.0
mask_file = Demo.png   # the PNGGeo driver's image
"""


@pytest.mark.parametrize("boundary,riemann", [("IDWBaseGFM", "HLL"), ("CutCell", "StegerWarming")])
def test_reference_signature_driver(oracle_mod, tmp_path, boundary, riemann):
    PIL = pytest.importorskip("PIL.Image")
    from ibtfo_amd import workloads as W
    assert os.path.exists(DRIVER), "tests/drivers/ref_sig_driver is not built (run __graft_entry__.build())"
    nx, ny, steps = 128, 64, 40
    PIL.fromarray(W.bodies_mask(nx, ny)).save(tmp_path / "Demo.png")  # uint8 x 4 -> RGBA
    (tmp_path / "case.inp").write_text(INP.format(nx=nx, ny=ny, steps=steps, boundary=boundary, riemann=riemann,
                                                  t_end=100.0, clip=0, adv=0))
    out = tmp_path / "out"
    out.mkdir()
    print("ref_sig_driver %s %s" % (boundary, riemann), flush=True)
    r = subprocess.run([DRIVER, "case.inp", str(out)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    XY = np.fromfile(out / "XY.bin")
    U = np.fromfile(out / "U.bin")
    log = np.fromfile(out / "log.bin").reshape(-1, 3)
    assert len(log) == steps
    O = oracle_mod
    c = O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=100.0, max_iter=steps,
               CFL_number=0.5, Scheme="MUSCL", Riemann=riemann, Limiter="VanLeer", Boundary=boundary,
               case="uniform", edge="inflow", quirks=0, extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    c.phi = XY[2::6].copy()  # the mask level set computed on the device (its own test pins it)
    assert np.abs(c.phi - O.mask_phi(c, W.bodies_mask(nx, ny))).max() <= 1e-15
    o = O.run_oracle(c, steps, root_mode=1)
    assert o["steps"] == steps
    assert nan_equal_bits(XY[5::6], o["XY"][5::6]) == 0
    assert nan_equal_bits(U, o["U_OLD"]) == 0
    assert nan_equal_bits(log[:, 0], o["dt"][:steps]) == 0
    assert nan_equal_bits(log[:, 1], o["t"][:steps]) == 0
    assert np.allclose(log[:, 2], o["mass"][:steps], rtol=c.M * c.N * 2.0 ** -53, atol=0)


def test_ibtfo_run_writes_reference_vtk(oracle_mod, tmp_path):
    """ibtfo_run --vtk: WriteData's files at the reference driver's cadence
    (every plot_per steps from iteration 0, and after the loop), the last one
    byte-identical to the reference format of the oracle's state."""
    import ibtfo_amd as ib
    exe = os.path.join(ROOT, "ibtfo_amd", "ibtfo_run")
    assert os.path.exists(exe), "ibtfo_amd/ibtfo_run is not built"
    O = oracle_mod
    c = O.Case(N_x=64, N_y=48, hi_x=1.0, hi_y=1.0, Boundary="IDWBaseGFM", case="incline", Psy_time=0.3,
               max_iter=23)
    inp = O.write_inp(c, str(tmp_path / "case.inp"))
    with open(inp, "a") as f:
        f.write("Case = incline\nplot_per = 10\n")
    vdir = tmp_path / "vtk"
    vdir.mkdir()
    print("ibtfo_run --vtk", flush=True)
    r = subprocess.run([exe, inp, "--quiet", "--chunk", "8", "--vtk", str(vdir)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (vdir / "result.visit").read_text() == "!NBLOCKS 1\ndat_0_0.vtk\ndat_0_10.vtk\ndat_0_20.vtk\ndat_0_23.vtk\n"
    o = O.run_oracle(c, 23, root_mode=1)
    ref = tmp_path / "ref"
    ref.mkdir()
    cfg = ib.config(N_x=64, N_y=48, hi_x=1.0, hi_y=1.0, gamma=1.4)
    ib.write_vtk_host(cfg, 23, float(o["t"][-1]), o["U_OLD"], o["XY"], ref)
    assert (vdir / "dat_0_23.vtk").read_bytes() == (ref / "dat_0_23.vtk").read_bytes()


def test_time_advance_takes_the_drivers_dt(oracle_mod, tmp_path):
    """TimeAdvance advances by the dt the driver passes (TimeAdvance.H:13-26,
    SingleStep.cpp:5-20): a driver that clips its last step to end at Psy_time
    gets the oracle's result of the same clipped loop, bit for bit; the same
    driver stepping through Advance writes the same files."""
    PIL = pytest.importorskip("PIL.Image")
    from ibtfo_amd import workloads as W
    assert os.path.exists(DRIVER), "tests/drivers/ref_sig_driver is not built (run __graft_entry__.build())"
    nx, ny = 128, 64
    mask = W.bodies_mask(nx, ny)
    PIL.fromarray(mask).save(tmp_path / "Demo.png")
    O = oracle_mod
    c = O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=100.0, max_iter=40,
               CFL_number=0.5, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM",
               case="uniform", edge="inflow", quirks=0, extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    c.phi = O.mask_phi(c, mask)
    st = O.setup(c)
    dt0 = O.compute_dt(c, st["U_OLD"], st["XY"])
    free = O.run_oracle(c, 40)
    # the driver's clock after n steps: dt0 + dt_1 + ... + dt_(n-1); end 40% into step 31
    t_end = float(dt0 + free["dt"][:29].sum() + 0.4 * free["dt"][29])
    c.Psy_time = t_end
    c.max_iter = 100
    o = O.run_oracle(c, 100, root_mode=1, clip=True)
    assert o["steps"] in (31, 32)
    (tmp_path / "case.inp").write_text(INP.format(nx=nx, ny=ny, steps=100, boundary="IDWBaseGFM", riemann="HLL",
                                                  t_end=t_end, clip=1, adv=0))
    out = tmp_path / "out"
    out.mkdir()
    print("ref_sig_driver clipped last step, t_end=%r" % t_end, flush=True)
    r = subprocess.run([DRIVER, "case.inp", str(out)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    U = np.fromfile(out / "U.bin")
    log = np.fromfile(out / "log.bin").reshape(-1, 3)
    assert len(log) == o["steps"]
    assert log[-1, 1] >= t_end and log[-2, 1] < t_end
    assert nan_equal_bits(U, o["U_OLD"]) == 0
    assert nan_equal_bits(log[:, 0], o["dt"]) == 0
    assert nan_equal_bits(log[:, 1], o["t"]) == 0
    # the unclipped loop of the same steps differs: the clipped dt was used
    f = O.run_oracle(c, o["steps"], root_mode=1)
    assert nan_equal_bits(f["U_OLD"], o["U_OLD"]) > 0
    # the same loop through Advance (Advance.H:14-26), which is what SingleStep's
    # TimeAdvance calls: the same files
    (tmp_path / "case.inp").write_text(INP.format(nx=nx, ny=ny, steps=100, boundary="IDWBaseGFM", riemann="HLL",
                                                  t_end=t_end, clip=1, adv=1))
    out2 = tmp_path / "out_advance"
    out2.mkdir()
    print("ref_sig_driver through Advance", flush=True)
    r = subprocess.run([DRIVER, "case.inp", str(out2)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (out2 / "U.bin").read_bytes() == (out / "U.bin").read_bytes()
    assert (out2 / "log.bin").read_bytes() == (out / "log.bin").read_bytes()
