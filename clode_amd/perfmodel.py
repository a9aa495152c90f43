"""Algorithmic operation counts per member-step, for roofline accounting.

A member-step is one accepted step of one ensemble member (the reference's
``step`` counter, clode/cpp/transient.cl:54-58), including the work of its
observer update.  Counting convention: add, sub, mul and div count 1 each, a
fused multiply-add counts 2 (it is written as a*b+c in the source);
comparisons, fmin/fmax/fabs, integer work and moves count 0;
transcendentals (exp, log, pow, sqrt) are counted separately.  Compile-time
constant expressions (tableau quotients, RCONST products) count 0.

The counts below are read off the device source expression by expression:
  steppers   clode_amd/csrc/device/clode_steppers.h
  observers  clode_amd/csrc/device/clode_observers.h
  models     clode_amd/models/*.cl
Rejected step attempts are not counted (the achieved rate is a lower bound).
"""
from __future__ import annotations

# right-hand sides: (flops, transcendentals) per evaluation
RHS = {
    "van_der_pol": (5, 0),          # y ; mu*(1 - x*x)*y - x
    "lorenz": (9, 0),               # s*(y-x) ; r*x - y - x*z ; x*y - b*z
    "lactotroph": (40, 3),          # 3 Boltzmann gates (4 + exp each), 5 currents, 4 derivatives
    "lactotroph_noise": (42, 3),    # + sigma*w and one more add
    "chay_keizer": (30, 2),         # 2 gates, omega with pown (3 mul), 3 currents, 3 derivatives
    "ornstein_uhlenbeck": (4, 0),   # theta*(mu - x) + sigma*w
    "harmonic": (3, 0),
    "decay": (2, 0),
    "sine": (2, 1),
    "linear_drag": (4, 0),
}

# stepper: (flops per variable, flops per step, RHS evaluations, transcendentals per step)
STEPPER = {
    "euler": (2, 1, 1, 0),          # x += dt*k1 ; t += dt
    "seuler": (2, 1, 1, 1),         # + Wiener draw: randn / sqrt(dt) (1 div per noise)
    "heun": (6, 1, 2, 0),           # fma stage ; x += dt*0.5*(k1+k2)
    "rk4": (14, 3, 4, 0),           # 3 stages x (mul+add), update 8 per var ; h2, th2, th
    "bs23": (22, 14, 3, 1),         # 2 stages, update, error (8), err/scale (1 div) ; times + controller ; pow
    "dopri5": (59, 20, 6, 1),       # 5 stages (35), update (11), error (12), err/scale (1) ; times + controller
}


def observer_flops(observer: str, n_var: int, n_aux: int, precision: str) -> int:
    """Per-step observer update cost (events are rare and not counted)."""
    if precision == "fp32c":
        # dt, then one double fma per variable/aux integral; downstate integral
        base = {"thresh2": 1 + 2, "localmax": 1, "nhood2": 1, "": 0}[observer]
        per = 2 * n_var + 2 * n_aux
        if observer == "nhood2":
            per += 4 * n_var  # distance to the neighbourhood centre
        return base + per if observer else 0
    if observer == "thresh2":
        # dt, elapsed ; running mean of dt (3) ; running time means (4 per var/aux) ; downstate mean (5)
        return 2 + 3 + 4 * n_var + 4 * n_aux + 5
    if observer == "localmax":
        return 2 + 4 * n_var + 4 * n_aux
    if observer == "nhood2":
        # + normalised distance (sub, div, mul, add per var) once the centre is found
        return 2 + 3 + 4 * n_var + 4 * n_aux + 4 * n_var
    return 0


def flops_per_step(model: str, stepper: str, observer: str, n_var: int, n_aux: int,
                   precision: str = "double", n_noise: int = 0) -> dict:
    rhs_f, rhs_t = RHS[model]
    per_var, per_step, n_rhs, trans = STEPPER[stepper]
    flops = n_rhs * rhs_f + per_var * n_var + per_step + observer_flops(observer, n_var, n_aux, precision)
    if stepper == "seuler":
        flops += n_noise  # division by sqrt(dt) per noise term
    return {"flops": flops, "transcendentals": n_rhs * rhs_t + trans + (2 * n_noise if stepper == "seuler" else 0)}


# MI355X published peaks (dense)
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16_mfma": 2500.0}
PEAK_HBM_GBS = 8000.0
