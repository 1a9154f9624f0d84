"""CPU checks of the dense torch reference used by the adjoint tests: its
forward step agrees with the oracle (the pinned CPU restatement), and its
autograd gradient agrees with central finite differences."""
import numpy as np
import torch

from oracle import oracle as O
from tests import adjoint_cases as AC
from tests import dense_reference as D

DT = 0.005


def _oracle_positions(arr, steps):
    a = arr.copy()
    sc = O.Scenario(cell_size=3.6, dt_min=DT, dt_max=DT, fluct=False, obstacle=False,
                    navigation=False)
    out = [a['position'].copy()]
    for k in range(steps):
        O.run(a, sc, 1, first_step=k)
        out.append(a['position'].copy())
    return np.stack(out)


def _params(arr):
    return (torch.tensor(3.0, dtype=torch.float64), torch.tensor(1.2e5, dtype=torch.float64),
            torch.tensor(4e4, dtype=torch.float64), torch.tensor(500.0, dtype=torch.float64),
            torch.as_tensor(arr['target_velocity'].copy(), dtype=torch.float64))


def test_dense_reference_matches_oracle():
    for arr in (AC.sparse(30, seed=3), AC.packed(25, seed=4)):
        steps = 25
        ref = _oracle_positions(arr, steps)
        with torch.no_grad():
            X = D.rollout(arr, steps, DT, 3.6, *_params(arr)).numpy()
        scale = np.abs(ref).max() + 1.0
        assert np.max(np.abs(X - ref)) < 1e-11 * scale


def test_packed_case_has_contacts_sparse_has_none():
    arr = AC.packed(25, seed=4)
    with torch.no_grad():
        X = D.rollout(arr, 10, DT, 3.6, *_params(arr))
    assert float(AC.min_skin(X, arr['radius']).max()) < 0.0
    arr = AC.sparse(30, seed=3)
    with torch.no_grad():
        X = D.rollout(arr, 10, DT, 3.6, *_params(arr))
    assert float(AC.min_skin(X, arr['radius']).min()) > 0.0


def test_dense_reference_gradient_matches_fd():
    arr = AC.packed(16, seed=5)
    steps = 12
    p = [t.clone().requires_grad_(True) for t in _params(arr)]
    q = [t.clone() for t in _params(arr)]
    q[0] = q[0] * 1.05
    q[2] = q[2] * 1.05
    q[4] = q[4] * 1.03
    with torch.no_grad():
        target = D.rollout(arr, steps, DT, 3.6, *q)

    def loss(params):
        X = D.rollout(arr, steps, DT, 3.6, *params)
        return ((X[1:] - target[1:]) ** 2).sum(-1).mean()

    L = loss(p)
    L.backward()
    for k in range(4):
        base = [t.detach().clone() for t in p]
        h = 1e-6 * max(1.0, abs(float(base[k])))
        up = [t.clone() for t in base]
        dn = [t.clone() for t in base]
        up[k] = up[k] + h
        dn[k] = dn[k] - h
        with torch.no_grad():
            fd = (loss(up) - loss(dn)) / (2 * h)
        g = float(p[k].grad)
        # central-difference truncation error plus its rounding floor eps L / h
        floor = 100 * np.finfo(np.float64).eps * float(L.detach()) / h
        assert abs(g - float(fd)) <= 1e-5 * abs(g) + floor, (k, g, float(fd), floor)
