"""C-ABI checks that need no GPU: the library loads, exports every declared entry point, and its
host-side helpers agree with the oracle bit for bit."""
import os
import re

import numpy as np
import pytest

import spp_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header=None):
    """Entry points declared in include/*.h (or one header)."""
    names = []
    for h in ([header] if header else sorted(os.listdir(os.path.join(ROOT, "include")))):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?(spp_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_what_binding_binds():
    from spp_amd import surrogate

    assert set(declared_functions("spp_engine.h")) == set(S.SIGNATURES)
    assert set(declared_functions("spp_surrogate.h")) == set(surrogate.SIGNATURES)


def test_library_exports_every_declared_symbol(engine_mod):
    L = engine_mod.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_params_default_and_abi_version(engine_mod):
    p = engine_mod.default_params(S.MODEL_REPULSIVE)
    assert p.abi_version == S.ABI_VERSION
    # reference globals and statics (shared/parameters.h:60-99, particle.h:774-786)
    assert p.Lx == 64 and p.dt == 1.0 / 1024 and p.r_list == 1.0625
    assert p.A_rep == 2 and p.sigma_rep == 1 and p.rc_rep == 1.2 and p.R_align == 1.1
    c = engine_mod.default_params(S.MODEL_CONTINUOUS)
    assert c.R_align == 1 and c.g == 4 and c.repulsion == S.REP_NONE


def test_create_rejects_bad_parameters(engine_mod):
    p = engine_mod.default_params(S.MODEL_REPULSIVE, n=100, Lx=2.0, Ly=2.0, r_list=1.5)
    with pytest.raises(S.SppError) as ei:
        S.Engine(p)  # two cells per side: rejected before any device call
    assert ei.value.code == S.ERR_ARG
    p = engine_mod.default_params(S.MODEL_REPULSIVE, n=100, r_list=0.5)
    with pytest.raises(S.SppError):
        S.Engine(p)  # list radius below the cutoff
    p = engine_mod.default_params(S.MODEL_REPULSIVE, n=100)
    p.abi_version = 99
    with pytest.raises(S.SppError):
        S.Engine(p)
    p = engine_mod.default_params(S.MODEL_REPULSIVE, n=100, kernel_variant=S.KERNEL_SLOT_DEEP + 1)
    with pytest.raises(S.SppError) as ei:
        S.Engine(p)
    assert ei.value.code == S.ERR_ARG
    for v in (1, 2):  # retired variants
        p = engine_mod.default_params(S.MODEL_REPULSIVE, n=100, kernel_variant=v)
        with pytest.raises(S.SppError) as ei:
            S.Engine(p)
        assert ei.value.code == S.ERR_UNSUPPORTED
    p = engine_mod.default_params(S.MODEL_ABP, n=100, kernel_variant=S.KERNEL_SLOT_DEEP)
    with pytest.raises(S.SppError) as ei:
        S.Engine(p)  # the active Brownian model runs on the slot kernel only
    assert ei.value.code == S.ERR_UNSUPPORTED


def test_host_noise_matches_oracle(engine_mod, oracle_mod):
    """Same Philox counters and Box-Muller transform; the engine's cosine (spp_amd/csrc/trig.hpp,
    at most 1.4 ulp) against the C library's in the oracle: equal to a few ulp."""
    for seed, i, step in [(1, 0, 0), (10, 12345, 678), (2**40 + 3, 262143, 9999), (7, 5, 2**33)]:
        a, b = S.noise_normal(seed, i, step), oracle_mod.noise_normal(seed, i, step)
        assert a == pytest.approx(b, rel=5e-16, abs=5e-16)
    z = np.array([S.noise_normal(3, i, 11) for i in range(2000)])
    w = np.array([oracle_mod.noise_normal(3, i, 11) for i in range(2000)])
    assert np.abs(z - w).max() <= 4.5e-16 * np.maximum(1.0, np.abs(w)).max()


def test_host_random_state_matches_oracle(engine_mod, oracle_mod):
    a = S.random_state(12, 1000, 32.0, 16.0)
    b = oracle_mod.random_state(12, 1000, 32.0, 16.0)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
