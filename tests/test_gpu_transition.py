"""Transition Hamiltonian (psoci_build_transition) against the oracle's
restatement of constructGlobalTransitionHamiltonian2WAY + packSefsef, bit
for bit: a list with itself (= the stored H's lower part), with a shifted
and a permuted copy of itself, with a list of configurations of other
jsym (spin-orbit blocks), and the refusal where the reference reads past a
one-entry diagonal vnt."""
import numpy as np
import pytest

import oracle
from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


def sub(cfg, occ, jsym):
    return inputs.Configs(nbf=cfg.nbf, nelec=cfg.nelec, ksym=cfg.ksym, occ=np.ascontiguousarray(occ, np.uint8),
                          jsym=np.ascontiguousarray(jsym, np.int32), title="list 2")


def class_permutation(cfg, seed):
    """Configurations with two or more open shells shuffled among those of
    their open-shell count, the others kept in place: equal positions then
    hold configurations with equal SEF counts and, where a double excitation
    apart, two or more open shells -- the inputs for which the reference's
    transition build stays within its arrays."""
    nopen = (cfg.occ == 1).sum(1)
    perm = np.arange(cfg.ncfg)
    rng = np.random.default_rng(seed)
    for s in np.unique(nopen):
        if s < 2:
            continue
        idx = np.where(nopen == s)[0]
        perm[idx] = idx[rng.permutation(len(idx))]
    return perm


def lists():
    cfg, ints, _ = inputs.workload("fci6")
    yield "self", cfg, ints, cfg
    p = class_permutation(cfg, 1)
    yield "permuted", cfg, ints, sub(cfg, cfg.occ[p], cfg.jsym[p])
    yield "longer", cfg, ints, sub(cfg, np.concatenate([cfg.occ[::-1], cfg.occ]), np.concatenate([cfg.jsym] * 2))
    p = class_permutation(cfg, 2)
    yield "prefix", cfg, ints, sub(cfg, np.concatenate([cfg.occ[:40], cfg.occ[p][40:100]]), cfg.jsym[:100])
    base = inputs.fci_configs(5, 3)
    ints5 = inputs.make_integrals(5, 5)
    yield "spin-orbit", base, ints5, sub(base, base.occ, np.full(base.ncfg, 3))
    base = inputs.fci_configs(4, 4)
    yield "spin-orbit-even", base, inputs.make_integrals(4, 6), sub(base, base.occ, np.full(base.ncfg, 3))


@pytest.mark.parametrize("case", [c[0] for c in lists()])
def test_transition_bit_identical(case):
    name, cfg, ints, cfg2 = next(c for c in lists() if c[0] == case)
    ref = oracle.OracleSpace(cfg, ints).transition_h(oracle.OracleSpace(cfg2, ints))
    H = api.Hamiltonian(cfg, ints)
    got = H.transition(cfg2)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b), name
    if name == "self":
        H.build()
        rowptr, col, val, _ = H.get_h()
        assert np.array_equal(got[0], rowptr) and np.array_equal(got[1], col) and np.array_equal(got[2], val)
    H.close()


@pytest.mark.parametrize("seed", [0, 1, 3])
def test_transition_refusal_matches_the_oracle(seed):
    """Random permutations put configurations with fewer SEFs (seed 0) or a
    closed shell a double excitation away (seeds 1, 3) at equal positions,
    where the reference reads out of bounds: both refuse."""
    cfg, ints, _ = inputs.workload("fci6")
    perm = np.random.default_rng(seed).permutation(cfg.ncfg)
    cfg2 = sub(cfg, cfg.occ[perm], cfg.jsym[perm])
    with pytest.raises(RuntimeError):
        oracle.OracleSpace(cfg, ints).transition_h(oracle.OracleSpace(cfg2, ints))
    H = api.Hamiltonian(cfg, ints)
    with pytest.raises(api.PsociError) as e:
        H.transition(cfg2)
    assert e.value.code == api.ERR_UNSUPPORTED
    with pytest.raises(api.PsociError) as e:
        H._check(H.L.psoci_get_transition(H._h, None, None, None))
    assert e.value.code == api.ERR_STATE
    H.close()
