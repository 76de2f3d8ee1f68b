"""Predicted output noise of the blind rotation from the parameter set.

The standard external-product noise analysis of TFHE gate bootstrapping
(Chillotti et al., "TFHE: Fast Fully Homomorphic Encryption over the Torus",
J. Cryptology 2020, Theorem 4.1 / Corollary 3.14), summed over the steps of
one rotation, for k = 1, binary keys and balanced gadget digits. One step is
ExtProd(d, C) with C a TRGSW (or a key-bundle / block combination of TRGSWs)
of message m; its noise is the digits times the key noise plus m times the
decomposition rounding error of d:

    Var = steps * [ 2 l N (Bg^2 - 1)/12 * key_terms * alpha^2     digits x key noise
                  + E|m|^2 (1 + N/2) Bg^(-2l) / 12 ]              rounding x message

key_terms: key-noise copies in one step's key: 1 for the plain CMUX
(acc += ExtProd((X^a - 1) acc, BK_i)), 6 for the 2-bit key bundle (three
TRGSWs, each times X^a - 1 = two monomials), 8 for block-binary blocks of
four. E|m|^2: the squared norm of the key's message: s_i in {0, 1} with
probability 1/2 each for the plain CMUX (1/2), a monomial for the key
bundle (1), X^a - 1 (two monomials) in a block that holds a key bit for the
block-binary key (2 x 4/5: blocks are drawn uniformly over their ell + 1 = 5
patterns). The output TLWE of a gate (after SampleExtract) carries this
variance in torus units squared; parameters are lvl1param's
(include/tfhe_params.hpp).
"""

LVL1_ALPHA = 3.42338787018369e-08  # 2^-24.8

KEY_TERMS = {"plain": 1, "kb": 6, "bb": 8}
MESSAGE_NORM = {"plain": 0.5, "kb": 1.0, "bb": 2.0 * 4 / 5}
STEPS = {"plain": 636, "kb": 318, "bb": 159}


def predicted_br_variance(steps, key_terms, message_norm, N=1024, l=2, bgbit=8,
                          alpha=LVL1_ALPHA):
    """Predicted output variance (torus units squared) of one rotation."""
    bg = 1 << bgbit
    digits = 2 * l * N * (bg * bg - 1) / 12.0 * key_terms * alpha * alpha
    rounding = message_norm * (1 + N / 2.0) * 2.0 ** (-2 * l * bgbit) / 12.0
    return steps * (digits + rounding)


def predicted_lvl1_variance(variant):
    """Predicted lvl1 gate output variance for variant 'plain', 'kb' or 'bb'."""
    return predicted_br_variance(STEPS[variant], KEY_TERMS[variant], MESSAGE_NORM[variant])
