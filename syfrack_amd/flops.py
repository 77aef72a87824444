"""Algorithmic operation counts of the distance-estimator path.

Convention: add, sub and mul count 1 FLOP, a fused multiply-add 2 (the library is built
without contraction, so none occur); division, square root and logarithm are special
operations counted apart; min, max, abs, clamp, select and compare are not counted.  The
counts are taken from the kernel arithmetic (syfrack_amd/csrc/device.hpp), expression by
expression, not from the ISA.

Units the kernels count on the device (one atomic per wave): Iterate-body executions
(``de_iterations``), field evaluations (``de_evals``), march steps (``steps``) and hits;
the FP64 refine counts its own evaluations and body executions.
"""
from __future__ import annotations

from . import abi

# FLOPs of one Iterate-body execution, per family (device.hpp de_eval):
#  Mandelbulb power 8: |z| 5, rho 3, angle squares 6 x 5 = 30, radial powers 5,
#     scaling 5, derivative 2, AddOrigin 4                          = 54  (+2 sqrt, 4 div)
#  Mandelbox: |z| 5, box fold 6, dot 5, sphere-fold scaling 4, Scale 4,
#     AddOrigin 4                                                    = 28  (+2 sqrt, <=1 div)
#  Menger: |z| 5, fold scaling 6, derivative 1, trap |z| 5            = 17  (+2 sqrt)
BODY = {abi.FAMILY_MANDELBULB: 54, abi.FAMILY_MANDELBOX: 28, abi.FAMILY_MENGER: 17, abi.FAMILY_SPHERE: 0}
# FLOPs of one evaluation outside the body: the final bailout |z| (5) plus the fence
# (bulb 0.5*ln(r)*r/dr: |z| 5 + 2; box |z|/|dr|: 5; menger box: 9; sphere graph 5 + 5).
EVAL = {abi.FAMILY_MANDELBULB: 12, abi.FAMILY_MANDELBOX: 10, abi.FAMILY_MENGER: 14, abi.FAMILY_SPHERE: 10}
# One march step outside the field: ro + rd*t 6, radius 1, over-relaxation test 2,
# step length 1, end test 1, advance 1 (+1 div for the relative error).
STEP = 12
# One hit's normal outside its four field evaluations: four taps p + k*h (6 each),
# gradient accumulation (6 each), normalisation 8 (+1 sqrt, 1 div).
NORMAL = 56
# One FP64 refine evaluation outside the field: ro + rd*t 6, |d|*safety - cone*t 3.
REFINE_EVAL = 9


def march_flops(family: int, stats: dict) -> float:
    """FP32 FLOPs of the primary trace's march (its evaluations and steps; the normal taps are
    counted apart, in hit_de_*)."""
    return (BODY[family] * (stats["de_iterations"] - stats["hit_de_iterations"])
            + EVAL[family] * (stats["de_evals"] - stats["hit_de_evals"]) + STEP * stats["steps"])


def hit_fp32_flops(family: int, stats: dict) -> float:
    """FP32 FLOPs of the primary trace's hit shading: the four-tap normals."""
    return BODY[family] * stats["hit_de_iterations"] + EVAL[family] * stats["hit_de_evals"] + NORMAL * stats["hits"]


def refine_flops(family: int, stats: dict) -> float:
    """FP64 FLOPs of the final-hit refines."""
    return (BODY[family] * stats["de64_iterations"] + (EVAL[family] + REFINE_EVAL) * stats["de64_evals"])


def tail_fp32_flops(family: int, stats: dict) -> float:
    """FP32 FLOPs of the normals shaded by the primary trace's trailing hit pass."""
    return (BODY[family] * stats["tail_hit_de_iterations"] + EVAL[family] * stats["tail_hit_de_evals"]
            + NORMAL * stats["tail_hits"])


def tail_fp64_flops(family: int, stats: dict) -> float:
    """FP64 FLOPs of the refines run by the primary trace's trailing hit pass."""
    return (BODY[family] * stats["tail_de64_iterations"]
            + (EVAL[family] + REFINE_EVAL) * stats["tail_de64_evals"])


def fused_flops(family: int, stats: dict) -> tuple:
    """(FP32, FP64) FLOPs of the fused march + hit-shading kernel alone: the march, and the
    normals and refines of the hits it shaded (all hits but the trailing pass's)."""
    fp32 = march_flops(family, stats) + hit_fp32_flops(family, stats) - tail_fp32_flops(family, stats)
    fp64 = refine_flops(family, stats) - tail_fp64_flops(family, stats)
    return fp32, fp64


def special_ops(family: int, stats: dict):
    """Special operations (correctly rounded division, square root, logarithm) of the fused
    march + hit-shading kernel alone, by type and precision, for the Mandelbulb families; None
    for the others, whose counts depend on branches the device does not count (the Mandelbox
    sphere fold divides only inside its shell).

    Per Iterate body: rho and |z| (2 sqrt), z / r, rho / r, x / rho, y / rho (4 div).  Per
    evaluation: the starting |z| (1 sqrt) and 0.5 * ln(r) * r / dr (1 log, 1 div).  Per march
    step: radius / t (1 div).  Per normal: the normalisation (1 sqrt, 1 div).  The FP64 refine
    evaluates the same field in FP64."""
    if family not in (abi.FAMILY_MANDELBULB,):
        return None
    bodies32 = stats["de_iterations"] - stats["tail_hit_de_iterations"]
    evals32 = stats["de_evals"] - stats["tail_hit_de_evals"]
    normals = stats["hits"] - stats["tail_hits"]
    bodies64 = stats["de64_iterations"] - stats["tail_de64_iterations"]
    evals64 = stats["de64_evals"] - stats["tail_de64_evals"]
    return {
        "fp32": {"div": 4 * bodies32 + evals32 + stats["steps"] + normals, "sqrt": 2 * bodies32 + evals32 + normals,
                 "log": evals32},
        "fp64": {"div": 4 * bodies64 + evals64, "sqrt": 2 * bodies64 + evals64, "log": evals64},
    }
