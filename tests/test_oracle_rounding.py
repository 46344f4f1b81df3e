"""Pins the oracle's rounding at the reference's @muladd sites.

The reference builds its smoothness indicators, WENO weighted sums and Centered interpolants as
metaprogrammed expressions (weno_interpolants.jl:228-265 `metaprogrammed_smoothness_operation`,
:267-283 the FP32 shifted variant, :492-516 `metaprogrammed_weno_reconstruction`;
reconstruction_coefficients.jl:135-165 `calc_reconstruction_stencil`, used under @muladd at
centered_reconstruction.jl:56-62).  This file rebuilds those expression trees from the same
generation rules, applies the MuladdMacro 0.2 rewrite (a sum of products evaluated left to right
becomes nested muladd calls, the first product rounded on its own and each later one fused into
the running sum; a product a·b·c splits as (a·b, c)), and evaluates them with exact rational
arithmetic rounded once per operation (a fused multiply-add rounds once).  The oracle must agree
bit for bit, and the cases are chosen where the unfused evaluation rounds differently, so that a
dropped or misplaced fusion fails.

What stays open (parity unpinned): MuladdMacro's source is not in the reference tree, so the
nesting above (left-to-right fold, (a·b, c) split) is stated, not read; nothing offline can
confirm it.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------------------------------------
# exact rounding
# ---------------------------------------------------------------------------------------------
def rd64(x: Fraction) -> float:
    return float(x)   # int / int true division is correctly rounded (half to even)


def rd32(x: Fraction) -> float:
    if x == 0:
        return 0.0
    s = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    m = round(x / Fraction(2) ** (e - 23))   # Fraction rounds half to even
    return s * float(m) * 2.0 ** (e - 23)


# expressions: floats are leaves, ("+", [...]) and ("*", [...]) are Julia's n-ary calls,
# ("-", a, b) a binary subtraction, ("fma", a, b, c) a muladd
def is_mul(e):
    return isinstance(e, tuple) and e[0] == "*"


def muladd(e):
    """The MuladdMacro rewrite of an expression (applied to every sub-expression): the sum's
    non-product terms (if any) start the running sum, otherwise its first product does; each
    remaining product, in order, is fused into the running sum."""
    if not isinstance(e, tuple):
        return e
    if e[0] == "-":
        return ("-", muladd(e[1]), muladd(e[2]))
    args = [muladd(a) for a in e[1]]
    if e[0] == "*":
        return (e[0], args)
    prods = [a for a in args if is_mul(a)]
    others = [a for a in args if not is_mul(a)]
    if others:
        acc, rest = (others[0] if len(others) == 1 else ("+", others)), prods
    else:
        acc, rest = prods[0], prods[1:]
    if not rest:
        return (e[0], args)
    for p in rest:
        f = p[1]
        x = f[0] if len(f) == 2 else ("*", f[:-1])
        acc = ("fma", x, f[-1], acc)
    return acc


def evaluate(e, rd):
    if not isinstance(e, tuple):
        return e
    if e[0] == "fma":
        a, b, c = (evaluate(x, rd) for x in e[1:])
        return rd(Fraction(a) * Fraction(b) + Fraction(c))
    if e[0] == "-":
        return rd(Fraction(evaluate(e[1], rd)) - Fraction(evaluate(e[2], rd)))
    vals = [evaluate(x, rd) for x in e[1]]
    acc = vals[0]
    for v in vals[1:]:
        acc = rd(Fraction(acc) * Fraction(v)) if e[0] == "*" else rd(Fraction(acc) + Fraction(v))
    return acc


# ---------------------------------------------------------------------------------------------
# the reference's generated expressions
# ---------------------------------------------------------------------------------------------
SMOOTHNESS = {3: [(10, -31, 11, 25, -19, 4), (4, -13, 5, 13, -13, 4), (4, -19, 11, 25, -31, 10)],
              2: [(1, -2, 1), (1, -2, 1)]}


def smoothness_expr(psi, C, buffer):
    """metaprogrammed_smoothness_operation(buffer): Σ_s ψ_s (Σ_{i≥s} C ψ_i) + ψ_b ψ_b C_last."""
    elem, c_idx = [], 1
    for st in range(1, buffer):
        terms = [("*", [float(C[c_idx + i - st - 1]), psi[i - 1]]) for i in range(st, buffer + 1)]
        elem.append(("*", [psi[st - 1], ("+", terms)]))
        c_idx += buffer - st + 1
    elem.append(("*", [psi[buffer - 1], psi[buffer - 1], float(C[c_idx - 1])]))
    return ("+", elem)


def stencils5(S, left):
    """S₀₃, S₁₃, S₂₃ of the six values ψ[i-3..i+2] (weno_interpolants.jl:451-453)."""
    s = [None] + list(S)   # 1-based like the reference
    if left:
        return [(s[3], s[4], s[5]), (s[2], s[3], s[4]), (s[1], s[2], s[3])]
    return [(s[4], s[3], s[2]), (s[5], s[4], s[3]), (s[6], s[5], s[4])]


def weno5_reference(S, left, fp32):
    rd = rd32 if fp32 else rd64
    cast = (lambda v: float(np.float32(v))) if fp32 else float
    S = [cast(v) for v in S]
    w5, _, _, _ = O.coefficients(fp32=fp32)
    Cstar = [cast(3 / 10), cast(3 / 5), cast(1 / 10)]
    eps = float(np.float32(1e-8))
    psis = stencils5(S, left)
    betas = []
    for r, psi in enumerate(psis):
        if fp32:   # ψ̂ = ψ[buffer ÷ 2 + 1] subtracted from every entry
            psi = [("-", v, psi[1]) for v in psi]
        betas.append(evaluate(muladd(smoothness_expr(list(psi), SMOOTHNESS[3][r], 3)), rd))
    tau = abs(rd(Fraction(betas[0]) - Fraction(betas[2])))
    alpha = []
    for r in range(3):
        q = rd(Fraction(tau) / Fraction(rd(Fraction(betas[r]) + Fraction(eps))))
        alpha.append(rd(Fraction(Cstar[r]) * Fraction(rd(1 + Fraction(rd(Fraction(q) * Fraction(q)))))))
    sigma = rd(Fraction(rd(Fraction(alpha[0]) + Fraction(alpha[1]))) + Fraction(alpha[2]))
    sinv = rd(1 / Fraction(sigma))
    omega = [rd(Fraction(a) * Fraction(sinv)) for a in alpha]
    # biased_p = sum(coeff .* ψ): element-wise products, summed left to right, no @muladd
    p = []
    for r, psi in enumerate(psis):
        prods = [rd(Fraction(float(w5[r, q])) * Fraction(psi[q])) for q in range(3)]
        p.append(rd(Fraction(rd(Fraction(prods[0]) + Fraction(prods[1]))) + Fraction(prods[2])))
    rec = ("+", [("*", [omega[r], p[r]]) for r in range(3)])
    return evaluate(muladd(rec), rd), betas


def unfused(e, rd):
    return evaluate(e, rd)


def offset_stencils(n, seed, offset, amp):
    rng = np.random.default_rng(seed)
    return [offset + amp * rng.uniform(-1, 1, 6) for _ in range(n)]


# ---------------------------------------------------------------------------------------------
# tests
# ---------------------------------------------------------------------------------------------
def test_muladd_rewrite_rule():
    """The assumed rewrite on small sums: a·b + c → muladd(a, b, c); c + a·b → muladd(a, b, c);
    a₁k₁ + a₂k₂ → muladd(a₂, k₂, a₁k₁) (Julia's left-to-right order kept); a·b·c + d →
    muladd(a·b, c, d)."""
    assert muladd(("+", [("*", ["a", "b"]), "c"])) == ("fma", "a", "b", "c")
    assert muladd(("+", ["c", ("*", ["a", "b"])])) == ("fma", "a", "b", "c")
    assert muladd(("+", [("*", ["a1", "k1"]), ("*", ["a2", "k2"])])) == ("fma", "a2", "k2", ("*", ["a1", "k1"]))
    assert muladd(("+", [("*", ["a", "b", "c"]), "d"])) == ("fma", ("*", ["a", "b"]), "c", "d")


@pytest.mark.parametrize("left", [True, False])
def test_weno5_smoothness_rounds_like_the_muladd_expression(left):
    differs = 0
    for S in offset_stencils(200, 11, 20.0, 1e-3):
        beta, _ = O.weno5_smoothness(S, left=left)
        for r, psi in enumerate(stencils5(list(S), left)):
            e = smoothness_expr(list(psi), SMOOTHNESS[3][r], 3)
            fused = evaluate(muladd(e), rd64)
            assert beta[r] == fused, (r, S, beta[r], fused)
            differs += fused != unfused(e, rd64)
    # the cases discriminate: many indicators round differently without the fusions
    assert differs > 50


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("left", [True, False])
def test_weno5_value_rounds_like_the_muladd_expressions(left, fp32):
    for seed, (off, amp) in enumerate([(20.0, 1e-3), (0.1, 1e-3), (0.0, 1.0), (-3.0, 1e-2)]):
        for S in offset_stencils(60, 100 + seed, off, amp):
            want, _ = weno5_reference(S, left, fp32)
            got = O.weno5(S, left=left, fp32=fp32)
            assert got == want, (S, got, want)


def test_weno3_value_rounds_like_the_muladd_expressions():
    w3 = O.coefficients()[1]
    rng = np.random.default_rng(5)
    differs = 0
    for _ in range(300):
        S = list(20.0 + 1e-3 * rng.uniform(-1, 1, 4))
        for left in (True, False):
            s = [None] + S
            psis = [(s[2], s[3]), (s[1], s[2])] if left else [(s[3], s[2]), (s[4], s[3])]
            betas = [evaluate(muladd(smoothness_expr(list(p), SMOOTHNESS[2][r], 2)), rd64) for r, p in enumerate(psis)]
            tau = abs(betas[0] - betas[1])
            eps = float(np.float32(1e-8))
            alpha = [c * (1 + (tau / (b + eps)) ** 2) for c, b in zip((2 / 3, 1 / 3), betas)]
            sinv = 1 / (alpha[0] + alpha[1])
            om_ = [a * sinv for a in alpha]
            p = [w3[r, 0] * psis[r][0] + w3[r, 1] * psis[r][1] for r in range(2)]
            rec = ("+", [("*", [om_[0], p[0]]), ("*", [om_[1], p[1]])])
            want = evaluate(muladd(rec), rd64)
            assert O.weno3(np.array(S), left=left) == want
            differs += want != unfused(rec, rd64)
    assert differs > 20


def test_centered4_rounds_like_the_muladd_expression():
    """calc_reconstruction_stencil(FT, 2, :symmetric) = c₄ψ[i-2] + c₃ψ[i-1] + c₂ψ[i] + c₁ψ[i+1]."""
    for fp32, rd in ((False, rd64), (True, rd32)):
        c4 = O.coefficients(fp32=fp32)[2]
        rng = np.random.default_rng(9)
        differs = 0
        for _ in range(400):
            S = 0.1 + 1e-3 * rng.uniform(-1, 1, 4)
            if fp32:
                S = S.astype(np.float32).astype(np.float64)
            e = ("+", [("*", [float(c4[q]), float(S[q])]) for q in range(4)])
            want = evaluate(muladd(e), rd)
            assert O.centered4(S, fp32=fp32) == want
            differs += want != unfused(e, rd)
        assert differs > 20
