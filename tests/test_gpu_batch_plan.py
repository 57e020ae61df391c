"""The batched solver's device round planner (k_batch.hip) against a host
restatement of its rules, on constructed round results.

The rules are those of the batched component solve (the per-component loop of
src/multi_component.rs:261-312 solved together): a segment stops when its
worst residual max_j ||L v_j - theta_j v_j|| / ||v_j|| is at or below tol,
when its Cholesky fails, or (FP32 rounds) after three rounds without a 0.7x
gain once below 1e-4; otherwise its damped interval becomes
[min(max(theta_m, 1.5 theta_k + 1e-12), 0.6), 2] and its filter degree the
largest d <= cap with T_d on that interval below e^9.2 at 0.  The planner's
gate is the largest degree over the iterating segments (0 past the round
limit), and its scalar table drives the segmented Chebyshev recurrence.
The test walks several planning calls in a row, so the carried state (stall
counters, best residuals, kept Ritz values) is checked across rounds too.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
si = pytest.importorskip("spectral_mi355x")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = si.Context(0)
    yield c
    c.close()


def degree(lo, cap):
    c, e = (lo + 2.0) / 2.0, (2.0 - lo) / 2.0
    x0 = max(c / e, 1.0 + 1e-12)
    return max(2, min(cap, int(math.floor(9.2 / math.acosh(x0)))))


def worst_of(norms2, c, k):
    wv = 0.0
    for j in range(k):
        r2, x2 = norms2[c, j], norms2[c, k + j]
        if math.isnan(r2) or math.isnan(x2):
            return math.inf
        xn = max(math.sqrt(max(x2, 0.0)), 2.220446049250313e-16)
        wv = max(wv, math.sqrt(max(r2, 0.0)) / xn)
    return wv


def host_plan(mode, update, st, status, theta, norms2, k, m, cap, steps, max_rounds, tol):
    """The planner's rules, restated on host copies of the state."""
    nseg = st["a"].size
    ctl = st["ctl"]
    ran = update == 2 or (update == 1 and ctl[0] > 0)
    deg = np.zeros(nseg, np.int32)
    for c in range(nseg):
        done = int(st["done"][c])
        if ran:
            if mode == 0:
                if done == 0:
                    if status[c] != 0:
                        done = 2
                    else:
                        wv = worst_of(norms2, c, k)
                        st["worst"][c] = wv
                        if wv <= tol:
                            done = 1
                        elif wv < 0.7 * st["best"][c]:
                            st["best"][c] = wv
                            st["stall"][c] = 0
                        else:
                            st["stall"][c] += 1
                            if st["stall"][c] >= 3 and wv < 1e-4:
                                done = 1
                        if done == 0:
                            lo = theta[c, k - 1] * 1.5 + 1e-12
                            st["a"][c] = min(max(theta[c, m - 1], lo), 0.6)
            elif done != 2:
                if status[c] != 0:
                    done = 2
                    st["worst"][c] = math.inf
                else:
                    st["thk"][c, :] = theta[c, :]
                    wv = worst_of(norms2, c, k)
                    st["worst"][c] = wv
                    if wv <= tol:
                        done = 1
            st["done"][c] = done
        if mode == 1 and done == 0:
            lo = st["thk"][c, k - 1] * 1.5 + 1e-12
            st["a"][c] = min(max(st["thk"][c, m - 1], lo), 0.6)
        deg[c] = 0 if done else degree(st["a"][c], min(cap, steps))
    gate = int(deg.max(initial=0))
    if ctl[1] >= max_rounds:
        gate = 0
    coef = np.zeros((steps, nseg + 1, 3))
    cast = np.float32 if mode == 0 else np.float64
    for j in range(1, steps + 1):
        for c in range(nseg + 1):
            d = deg[c] if c < nseg else 0
            if j <= d:
                lo = st["a"][c]
                cc, ee = (lo + 2.0) / 2.0, (2.0 - lo) / 2.0
                q = (1.0 / ee, -cc / ee, 0.0) if j == 1 else (2.0 / ee, -2.0 * cc / ee, -1.0)
            else:
                q = (0.0, 1.0, 0.0)
            coef[j - 1, c] = [float(cast(v)) for v in q]
    active = int(np.count_nonzero(deg))
    ctl[0] = gate
    if gate > 0:
        # an FP32 round after the first starts from the known L X: one SpMM fewer
        ctl[2] += gate + 1 - (1 if mode == 0 and ctl[1] > 0 else 0)
        ctl[1] += 1
    ctl[3] = active
    return deg, coef


def new_state(nseg, m, mode):
    st = dict(a=np.full(nseg, 0.1), best=np.full(nseg, np.inf), worst=np.full(nseg, np.inf),
              stall=np.zeros(nseg, np.int32), done=np.zeros(nseg, np.int32),
              ctl=np.zeros(4, np.int32))
    if mode == 1:
        st["thk"] = np.zeros((nseg, m))
    return st


def copy_state(st):
    return {key: v.copy() for key, v in st.items()}


def round_results(rng, nseg, k, m, scale):
    """Ritz values (ascending), squared residual / norm sums whose worst column
    has the per-segment residual `scale` (the others 1/2, 1/4, ... of it), and a
    Cholesky status of 0."""
    theta = np.sort(rng.uniform(0.0, 1.2, size=(nseg, m)), axis=1)
    theta[:, 0] = rng.uniform(0, 1e-9, size=nseg)
    x2 = rng.uniform(0.5, 2.0, size=(nseg, k))
    frac = 0.5 ** np.arange(k)
    r2 = (np.asarray(scale)[:, None] * frac[None, :]) ** 2 * x2
    return np.zeros(nseg, np.int32), theta, np.concatenate([r2, x2], axis=1)


def check(dev_st, host_st, deg_d, coef_d, deg_h, coef_h):
    for key in host_st:
        np.testing.assert_array_equal(dev_st[key], host_st[key], err_msg=key)
    np.testing.assert_array_equal(deg_d, deg_h)
    np.testing.assert_array_equal(coef_d, coef_h)


@pytest.mark.parametrize("cap", [24, 9])
def test_planner_fp32_rounds_match_host_rules(ctx, cap):
    print(f"case planner FP32 rounds cap={cap}", flush=True)
    rng = np.random.default_rng(7 + cap)
    nseg, k, m, steps, tol, max_rounds = 7, 3, 16, cap, 1e-6, 5
    dev, host = new_state(nseg, m, 0), new_state(nseg, m, 0)
    deg_d, coef_d = si.batch_plan(ctx, 0, 0, dev, *round_results(rng, nseg, k, m, np.ones(nseg)),
                                  k, m, cap, steps, max_rounds, tol)
    deg_h, coef_h = host_plan(0, 0, host, *round_results(np.random.default_rng(7 + cap), nseg, k,
                                                           m, np.ones(nseg)),
                              k, m, cap, steps, max_rounds, tol)
    check(dev, host, deg_d, coef_d, deg_h, coef_h)
    assert dev["ctl"][0] == degree(0.1, cap) and dev["ctl"][3] == nseg
    # residual levels per round: segment 0 converges in round 2, segment 1 stalls
    # below 1e-4 (three rounds without a 0.7x gain), segment 2 fails its
    # Cholesky in round 3, segment 3 gets a NaN norm, segments 4-6 keep improving
    levels = [
        [1e-2, 5e-5, 1e-1, 1e-1, 1e-1, 1e-2, 1e-3],
        [5e-7, 4.9e-5, 1e-2, 1e-2, 1e-2, 1e-3, 1e-4],
        [5e-7, 4.8e-5, 1e-3, np.nan, 1e-3, 1e-4, 1e-5],
        [5e-7, 4.7e-5, 1e-4, 1e-3, 1e-4, 1e-5, 2e-6],
        [5e-7, 4.6e-5, 1e-5, 1e-4, 1e-5, 9e-7, 1e-7],
    ]
    for r, lev in enumerate(levels):
        res_rng = np.random.default_rng(100 + r)
        status, theta, norms2 = round_results(res_rng, nseg, k, m, np.nan_to_num(lev, nan=1.0))
        if np.isnan(lev[3]):
            norms2[3, 1] = np.nan
        if r == 2:
            status[2] = 1
        deg_d, coef_d = si.batch_plan(ctx, 0, 1, dev, status, theta, norms2, k, m, cap, steps,
                                      max_rounds, tol)
        deg_h, coef_h = host_plan(0, 1, host, status, theta, norms2, k, m, cap, steps,
                                  max_rounds, tol)
        check(dev, host, deg_d, coef_d, deg_h, coef_h)
        if r == 2:
            assert dev["worst"][3] == np.inf and dev["done"][3] == 0  # NaN: not converged
    assert list(dev["done"]) == [1, 1, 2, 0, 0, 1, 1]
    assert dev["stall"][1] == 3
    assert dev["ctl"][1] == max_rounds and dev["ctl"][0] == 0  # the round limit closes the gate
    assert dev["ctl"][3] == 2


def test_planner_skips_results_of_an_idle_round(ctx):
    print("case planner idle round", flush=True)
    rng = np.random.default_rng(3)
    nseg, k, m = 4, 3, 16
    st = new_state(nseg, m, 0)
    st["done"][:] = 1
    st["ctl"][:] = [0, 2, 30, 0]  # the last gate was 0: nothing ran
    before = copy_state(st)
    status, theta, norms2 = round_results(rng, nseg, k, m, np.full(nseg, 1.0))
    status[:] = 1
    deg, coef = si.batch_plan(ctx, 0, 1, st, status, theta, norms2, k, m, 24, 24, 60, 1e-6)
    for key in ("a", "best", "worst", "stall", "done"):
        np.testing.assert_array_equal(st[key], before[key], err_msg=key)
    assert not deg.any() and list(st["ctl"]) == [0, 2, 30, 0]
    np.testing.assert_array_equal(coef[:, :, 0], 0.0)
    np.testing.assert_array_equal(coef[:, :, 1], 1.0)


def test_planner_fp64_rounds_match_host_rules(ctx):
    print("case planner FP64 rounds", flush=True)
    nseg, k, m, cap, steps, tol = 6, 3, 8, 24, 24, 1e-6
    dev, host = new_state(nseg, m, 1), new_state(nseg, m, 1)
    levels = [
        [1e-7, 1e-5, 1e-3, 1e-5, 1e-4, 1e-2],
        [2e-6, 1e-7, 1e-4, 1e-6, 1e-5, 1e-3],   # a converged segment keeps being re-read
        [3e-7, 5e-8, 1e-5, 1e-7, 1e-6, 1e-4],
    ]
    for r, lev in enumerate(levels):
        status, theta, norms2 = round_results(np.random.default_rng(200 + r), nseg, k, m, lev)
        if r == 1:
            status[5] = 1  # fails: its kept Ritz values stay, worst becomes inf
        upd = 2 if r == 0 else 1
        deg_d, coef_d = si.batch_plan(ctx, 1, upd, dev, status, theta, norms2, k, m, cap, steps,
                                      20, tol)
        deg_h, coef_h = host_plan(1, upd, host, status, theta, norms2, k, m, cap, steps, 20, tol)
        check(dev, host, deg_d, coef_d, deg_h, coef_h)
    assert dev["done"][5] == 2 and dev["worst"][5] == np.inf
    assert dev["done"][0] == 1 and dev["worst"][0] == host["worst"][0]
