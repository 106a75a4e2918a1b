"""CPU-side checks of the C ABI: the library loads, exports every entry point that
include/lbm_hip.h declares, and rejects invalid arguments before touching the GPU (the
reference's std::invalid_argument cases, src/v3_cuda_opt/solver_v3.cu:137-153,169-170)."""
from __future__ import annotations

import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_entry_points():
    text = (ROOT / "include" / "lbm_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point(lbm):
    L = lbm.lib()
    names = declared_entry_points()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_compiles_as_c(tmp_path):
    import shutil
    import subprocess

    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    src = tmp_path / "t.c"
    src.write_text('#include "lbm_hip.h"\nint main(void){lbm_solver* s=0; (void)s; return 0;}\n')
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", "-c", str(src),
                    "-o", str(tmp_path / "t.o")], check=True)


@pytest.mark.parametrize("nx,ny,tau,prec,msg", [
    (0, 10, 0.8, 0, "grid too small"),
    (10, 10, 0.5, 0, "tau must be > 0.5"),
    (10, 10, 0.4, 1, "tau must be > 0.5"),
    (10, 10, 0.8, 7, "precision"),
])
def test_create_rejects_invalid_arguments_without_gpu(lbm, nx, ny, tau, prec, msg):
    solid = np.zeros(max(nx * ny, 1), np.uint8)
    h = C.c_void_p()
    rc = lbm.lib().lbm_create(nx, ny, tau, 0.0, 0.0, solid, prec, C.byref(h))
    assert rc == lbm.LBM_EINVAL
    assert msg in lbm.last_error()
    assert not h.value


def test_python_mirror_raises_value_error(lbm):
    with pytest.raises(ValueError, match="solid mask size"):
        lbm.Solver(4, 4, 0.8, 0.0, 0.0, np.zeros(3, np.uint8))
    with pytest.raises(ValueError, match="tau must be > 0.5"):
        lbm.Solver(4, 4, 0.5, 0.0, 0.0, np.zeros(16, np.uint8))
    with pytest.raises(ValueError, match="unknown precision"):
        lbm.Solver(4, 4, 0.8, 0.0, 0.0, np.zeros(16, np.uint8), "bf16")


def test_null_handles_are_rejected(lbm):
    L = lbm.lib()
    assert L.lbm_step(None) == lbm.LBM_EINVAL
    assert L.lbm_run(None, 1) == lbm.LBM_EINVAL
    assert L.lbm_vorticity(None, None, 1) == lbm.LBM_EINVAL
    assert L.lbm_macroscopic(None, None, None, None) == lbm.LBM_EINVAL
    assert L.lbm_destroy(None) == lbm.LBM_OK


def test_poiseuille_force_matches_reference_formula(lbm):
    # real(8) * ((tau - 1/2) / 3) * u_max / (H * H) in the solver real (geometry.hpp:24-26)
    f32 = np.float32
    for H, tau in ((16, 1.0), (32, 0.5 + 3 ** 0.5 / 4), (64, 0.6)):
        ref = f32(8) * ((f32(tau) - f32(0.5)) / f32(3)) * f32(0.05) / (f32(H) * f32(H))
        assert lbm.poiseuille_force(H, tau, 0.05, "float") == float(ref)
        refd = 8.0 * ((tau - 0.5) / 3.0) * 0.05 / (float(H) * float(H))
        assert lbm.poiseuille_force(H, tau, 0.05, "double") == refd


def test_geometry_helpers_match_oracle(lbm, oracle_mod):
    for nx, ny in ((53, 31), (120, 48), (7, 3)):
        np.testing.assert_array_equal(lbm.make_channel(nx, ny), oracle_mod.make_channel(nx, ny))
    for args in ((120, 48, 30.0, 25.0, 10.0), (1024, 256, 256.0, 129.0, 32.0),
                 (512, 128, 128.0, 64.0, 24.0)):
        np.testing.assert_array_equal(lbm.make_cylinder(*args), oracle_mod.make_cylinder(*args))


def test_strouhal_recovers_a_known_frequency(lbm):
    """A pure tone of known frequency (period 371.3 steps, sampled every 50): both estimators
    must find it (the FFT with parabolic refinement to 1e-3, zero crossings to 1e-3)."""
    dt, period, D, U = 50.0, 371.3, 32.0, 0.05
    t = np.arange(4000) * dt
    uy = 0.01 * np.sin(2 * np.pi * t / period + 0.3) + 1e-4
    st_fft, st_zero, n = lbm.strouhal(uy, dt, D, U)
    exact = D / (period * U)
    assert abs(st_fft - exact) / exact < 1e-3
    assert abs(st_zero - exact) / exact < 1e-3
    assert n > 100


def test_seeded_block_matches_the_fixture_generator(lbm):
    """The package's seeded obstacle is the one the committed fixtures were written with."""
    import importlib.util
    from conftest import ROOT

    spec = importlib.util.spec_from_file_location("make_golden", ROOT / "tests" / "golden" / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for seed in (0, 1, 7, 123):
        assert lbm.seeded_block_position(seed, 53, 31) == mg.seeded_block_position(seed)
        assert (lbm.seeded_block_mask(seed, 53, 31) == mg.seeded_block_mask(seed)).all()
