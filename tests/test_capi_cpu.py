"""CPU-side checks of the C ABI library: it loads, exports exactly what
include/negf_capi.h declares, and refuses to run without a GPU (no host fallback).
Also pins the host-side Hamiltonian builder."""
import ctypes
import os
import re

import numpy as np
import pytest

from nonequilibriumgreensfunction_amd import capi
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire, splitmix64_uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "negf_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(negf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_exactly_the_exports():
    assert header_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    for name in header_functions():
        assert hasattr(lib, name), name


def _has_gpu():
    try:
        return capi.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a GPU is visible")
def test_no_device_fails_loudly():
    with pytest.raises(capi.NegfError) as e:
        capi.Context(4, 2, 2, 4)
    assert e.value.code == capi.NEGF_ERR_NO_DEVICE
    with pytest.raises(capi.NegfError):
        capi.invert_matrix(np.eye(4))


def test_bad_config_rejected_before_device():
    cfg = capi.NegfConfig(4, 2, 2, 3, 1e-6, 1e-10, 200, 1, 0, 0, 0.0, 0)  # norb = 3
    h = ctypes.c_void_p()
    assert capi.lib().negf_init(ctypes.byref(h), ctypes.byref(cfg)) == capi.NEGF_ERR_ARG


def test_server_lanes_out_of_range_rejected_before_device():
    for lanes in (-1, 5):
        cfg = capi.NegfServerConfig(1e-6, 1e-10, 200, 1, 0, 0, 0, 0.0, 0.0, lanes)
        h = ctypes.c_void_p()
        assert capi.lib().negf_server_create(ctypes.byref(h), ctypes.byref(cfg)) == capi.NEGF_ERR_ARG
        assert "lanes" in capi.lib().negf_last_error().decode()


def test_bench_gemm_bad_shape_rejected_before_device():
    ms = ctypes.c_double(0.0)
    for n, batch, reps in ((0, 1, 1), (64, 0, 1), (64, 1, 0), (9000, 1, 1)):
        assert capi.lib().negf_bench_gemm(0, 1, n, batch, reps, 1, ctypes.byref(ms)) == capi.NEGF_ERR_ARG


def test_splitmix64_stream_is_pinned():
    # first outputs of SplitMix64 seeded with 1 (published reference values of the generator)
    u = splitmix64_uniform(1, 3)
    expect = np.array([0x910A2DEC89025CC1, 0xBEEB8DA1658EEC67, 0xF893A2EEFB32555E], dtype=np.uint64)
    assert np.array_equal(u, (expect >> np.uint64(11)).astype(float) / 2.0**53)


def test_wire_structure():
    w = build_wire(6, 3, 2, disorder=0.7, seed=4, flux=0.3, zeeman=0.02)
    n, L, no = w.n, w.npx, w.norb
    assert w.htotd.shape == (n, n, L) and w.htotd.flags.f_contiguous
    for x in range(L):
        assert np.allclose(w.htotd[:, :, x], w.htotd[:, :, x].conj().T)  # Hermitian slices
    mask = np.kron(np.eye(n // no), np.ones((no, no))).astype(bool)
    assert np.all(w.hopx[~mask] == 0)  # site-block-diagonal HopX
    # disorder only on slices 2..Npx-1: end slices equal the clean slice
    clean = build_wire(6, 3, 2, flux=0.3, zeeman=0.02)
    assert np.array_equal(w.htotd[:, :, 0], clean.htotd[:, :, 0])
    assert np.array_equal(w.htotd[:, :, -1], clean.htotd[:, :, -1])
    assert not np.array_equal(w.htotd[:, :, 2], clean.htotd[:, :, 2])
    d = np.diag(w.htotd[:, :, 2] - clean.htotd[:, :, 2]).real
    assert np.all(np.abs(d) <= 0.35 + 1e-12)
    # Anderson term is a scalar per site: equal across the site's orbitals
    assert np.allclose(d.reshape(-1, no), d.reshape(-1, no)[:, :1])
