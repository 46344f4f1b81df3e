"""FP32 mode of the headline free-convection LES from its specified initial state (at rest,
b = N²z plus 1e-3·N²Δz noise in the top 16 m, no velocity noise), 10 steps of Δt = 1 s, against
the FP64 CPU oracle on the same FP32-rounded inputs.

Conditioning.  From rest, the velocities after 10 steps (~1e-8 m/s) are driven by buoyancy
anomalies of ~1e-8 on a background of up to 1.3e-3, i.e. at a few FP32 ulps of the field they
sit in, and FP64 itself moves by ~8e-4 when its inputs are perturbed at FP32 epsilon.  The
reference's own algorithm evaluated in FP32 (the oracle's FP32 mode, same inputs) lands 6e-3 from
FP64 at 16³ and 0.15 at 256×256×128: it forms the horizontal gradient of the hydrostatic pressure
anomaly as the difference of two column sums of b that agree to a few ulps.  The HIP FP32 path
forms that gradient as column sums of b's horizontal differences (exact in FP32), which brings
its velocities to ~1e-4 of FP64 at both sizes, below the problem's own sensitivity to FP32 input
rounding; the 1e-5 bar holds for the buoyancy, which is well conditioned, and the velocity bars
are the measured ones.
"""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import convection, oracle_for, product_for, rel_err

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "b")


def _state(model):
    return [(model.velocities[n] if q < 3 else model.tracers[n]).interior() for q, n in enumerate(NAMES)]


def test_fp32_from_rest_16cubed_against_oracle():
    grid32, kw, init = convection(N=(16, 16, 16), float_type=np.float32)
    grid64, kw64, _ = convection(N=(16, 16, 16), float_type=np.float64)
    model = product_for(grid32, kw, init)
    o64 = oracle_for(grid64, kw64, init, round_inputs_to=np.float32)
    o32 = oracle_for(grid32, kw, init, oracle_float_type=np.float32)
    model.time_step(1.0, 10)
    o64.time_step(1.0, 10)
    o32.time_step(1.0, 10)
    gpu = {n: rel_err(a, o64.get(q)) for q, (n, a) in enumerate(zip(NAMES, _state(model)))}
    ref32 = {n: rel_err(o32.get(q), o64.get(q)) for q, n in enumerate(NAMES)}
    print("16^3 from rest, 10 steps: HIP FP32 vs FP64 oracle", {k: f"{v:.2e}" for k, v in gpu.items()},
          "| oracle FP32 vs FP64 oracle", {k: f"{v:.2e}" for k, v in ref32.items()}, flush=True)
    assert gpu["b"] <= 1e-5
    for n in ("u", "v", "w"):
        # measured 8.1e-5, 7.9e-5, 1.4e-4 against the reference algorithm's 6.5e-3, 6.2e-3, 4.6e-3
        assert gpu[n] <= 3e-4 and gpu[n] <= ref32[n] / 10, (n, gpu[n], ref32[n])
    model.close()


# measured at 256×256×128 (FP32 HIP path, 10 steps from rest, max|a − b| / max|b| against the FP64
# oracle): velocities 8.4e-5 (u), 8.8e-5 (v), 1.5e-4 (w); the reference algorithm in FP32 (the
# oracle's FP32 mode) is 0.15–0.16 off, from the horizontal gradient of pHY′ taken as a difference of
# two rounded column sums, which the HIP path forms as column sums of buoyancy differences instead
PLAIN_FP32_VELOCITY_BOUND = 3e-4


def test_fp32_from_rest_full_size_against_oracle():
    """256×256×128 (the benchmark configuration), 10 steps, against the FP64 oracle and the oracle's
    FP32 mode on the same FP32-rounded inputs (each about a minute of the host's cores).  The
    oracle's FP32 mode evaluates the tendencies in FP32 but its Poisson solve in FP64; the HIP path
    with one FP64 iterative-refinement step per Poisson solve is the matching comparison and must
    stay within a factor 2 of it.  The plain FP32 HIP path (FP32 FFTs, as the benchmark runs) is held
    to its measured bound, which is more than a hundred times closer to FP64 than the reference
    algorithm's FP32 evaluation; the buoyancy to 1e-5."""
    N = (256, 256, 128)
    grid32, kw, init = convection(N=N, float_type=np.float32)
    grid64, kw64, _ = convection(N=N, float_type=np.float64)
    runs = {}
    for label, refine in (("plain", False), ("refined", True)):
        model = product_for(grid32, kw, init)
        model.set_tuning(poisson_refinement=refine)
        model.time_step(1.0, 10)
        runs[label] = _state(model)
        model.close()
    o64 = oracle_for(grid64, kw64, init, round_inputs_to=np.float32)
    o64.time_step(1.0, 10)
    ref64 = [o64.get(q) for q in range(4)]
    del o64
    errs = {label: {n: rel_err(a, ref64[q]) for q, (n, a) in enumerate(zip(NAMES, st))} for label, st in runs.items()}
    for label, e in errs.items():
        print(f"256x256x128 from rest, 10 steps: HIP FP32 ({label} Poisson) vs FP64 oracle",
              {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    o32 = oracle_for(grid32, kw, init, oracle_float_type=np.float32)
    o32.time_step(1.0, 10)
    ref32 = {n: rel_err(o32.get(q), ref64[q]) for q, n in enumerate(NAMES)}
    print("256x256x128 from rest, 10 steps: oracle FP32 (FP64 Poisson) vs FP64 oracle",
          {k: f"{v:.2e}" for k, v in ref32.items()}, flush=True)
    for label in errs:
        assert errs[label]["b"] <= 1e-5
    for n in ("u", "v", "w"):
        assert errs["refined"][n] <= 2 * ref32[n], (n, errs["refined"][n], ref32[n])
        assert errs["plain"][n] <= PLAIN_FP32_VELOCITY_BOUND, (n, errs["plain"][n])
        assert errs["plain"][n] <= ref32[n] / 100, (n, errs["plain"][n], ref32[n])
