"""Time-to-target estimators (reference gpu/onsager_v2_20261007/src/analyze.py:15-47, PROTOCOL.md section 6).

primary   : P_batch = fraction of batches with >= 1 success; TTS99 = t_batch ln(0.01) / ln(1 - P_batch) (t_batch if P = 1)
secondary : p = per-anneal success fraction; TTS99 = t_batch ln(0.01) / (B ln(1 - p)) (no floor)
intervals : Wilson 95 % on P_batch and p, mapped through the monotone TTS formulas.
"""
import math

Z95 = 1.959963984540054
LN01 = math.log(0.01)


def wilson(k, n, z=Z95):
    if n == 0:
        return (0.0, 1.0)
    ph = k / n
    den = 1 + z * z / n
    cen = (ph + z * z / (2 * n)) / den
    half = z * math.sqrt(ph * (1 - ph) / n + z * z / (4 * n * n)) / den
    return (max(0.0, cen - half), min(1.0, cen + half))


def tts_primary(P, t):
    if P >= 1.0:
        return t
    if P <= 0.0:
        return math.inf
    return t * LN01 / math.log(1.0 - P)


def tts_secondary(p, B, t):
    if p <= 0.0:
        return math.inf
    if p >= 1.0:
        return 0.0
    return t * LN01 / (B * math.log(1.0 - p))


def estimate(successes_per_batch, B, t_batch):
    """successes_per_batch: list of success counts (one per batch of B anneals); t_batch: mean batch time."""
    nb = len(successes_per_batch)
    k_b = sum(1 for s in successes_per_batch if s > 0)
    k_t = sum(successes_per_batch)
    Pb, p = k_b / nb, k_t / (B * nb)
    Pb_lo, Pb_hi = wilson(k_b, nb)
    p_lo, p_hi = wilson(k_t, B * nb)
    return dict(batches=nb, B=B, successes=k_t, p=p, p_wilson95=[p_lo, p_hi], P_batch=Pb, P_batch_wilson95=[Pb_lo, Pb_hi],
                t_batch=t_batch, tts_primary=tts_primary(Pb, t_batch),
                tts_primary_ci=[tts_primary(Pb_hi, t_batch), tts_primary(Pb_lo, t_batch)],
                tts_secondary=tts_secondary(p, B, t_batch),
                tts_secondary_ci=[tts_secondary(p_hi, B, t_batch), tts_secondary(p_lo, B, t_batch)])


def cut_from_sum_sh(sumw, sum_sh):
    """cut = (sumw + sum_sh / 2) / 2 with C integer division (sca_ref.hpp:154)."""
    def cdiv(a, b):
        q = abs(a) // abs(b)
        return q if (a >= 0) == (b > 0) else -q
    return cdiv(int(sumw) + cdiv(int(sum_sh), 2), 2)


def energy_target_cut(sumw, n, esu=1.499):
    """Pre-declared target: cut reached at energy density E[s u] (the K2000 33,000 mapping: cut = (sumw + N^1.5 E/2)/2)."""
    return int(math.ceil((sumw + 0.5 * n ** 1.5 * esu) / 2))
