"""CPU checks of the workload helpers in aar_hip.problems: the FD weights,
the spectral (exact) solution used to measure solver error, and the
uneven batch's instance list."""
import numpy as np

from aar_hip import problems as P


def test_lap_weights_match_oracle(oracle):
    """Bit-identical to the oracle's restatement of system.c:480-498 (pinned
    there by the closed forms and the reference build), for every order."""
    for fdn in range(1, 9):
        assert np.array_equal(np.array(P.lap_weights(fdn)), oracle.lap_coefficient(fdn)), fdn
    c = P.lap_weights(6)
    assert abs(c[0] + 2 * sum(c[1:])) < 1e-15           # zero row sum
    assert c[1] == 12.0 / 7.0 and c[6] == -1.0 / 16632.0


def test_spectral_solve_solves_the_discrete_system(oracle):
    """spectral_solve_torch (on the CPU here) returns u with a*Lap u + sigma
    u = b to rounding, through the oracle's stencil; for Poisson u has zero
    mean."""
    import torch
    shape = (12, 10, 8)
    for sigma, mean0 in ((0.0, True), (0.5, False)):
        b = P.rhs_normal(shape, 55)
        if mean0:
            b -= b.mean()
        u = P.spectral_solve_torch(torch.from_numpy(b), shape, -4.0, sigma).numpy()
        r = b - oracle.matvec(u, shape, -4.0, sigma)
        assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(b)
        if mean0:
            assert abs(u.mean()) < 1e-15 * np.abs(u).max() * u.size


def test_batch_instances():
    inst = P.batch_instances(1004)
    assert len(inst) == 112
    sides = {sh[0] for sh, _, _ in inst[:96]}
    assert sides <= {24, 32, 48, 64, 96, 128} and len(sides) == 6
    assert all(0.1 <= s <= 2.0 for _, s, _ in inst)
    assert all(sh == (128, 128, 128) for sh, _, _ in inst[96:])
    assert len({s for _, s, _ in inst[96:]}) == 1
    assert [k for _, _, k in inst] == list(range(100, 212))
