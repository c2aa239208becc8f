"""Codon space of the universal genetic code, in codeml's conventions.

Nucleotides are ordered T, C, A, G (tools.c:14) and a codon's 64-index is
16*b0 + 4*b1 + b2.  The 61 sense codons keep that order and are numbered
0..60 (FROM61/FROM64, treesub.c:2225-2268).  Ambiguous nucleotide letters
resolve as in tools.c:14-16; an ambiguous codon maps to the set of sense
codons it can stand for (SetMapAmbiguity, treesub.c:1146-1202).
"""
import numpy as np

BASES = "TCAG"
# amino acid of each of the 64 codons (TCAG order), '*' = stop; universal code
AA64 = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"
SENSE64 = [i for i in range(64) if AA64[i] != "*"]
NSENSE = len(SENSE64)                      # 61
FROM61 = np.array(SENSE64, dtype=np.int64)  # sense index -> 64-index
FROM64 = np.full(64, -1, dtype=np.int64)   # 64-index -> sense index or -1
FROM64[FROM61] = np.arange(NSENSE)
AA61 = np.array([AA64[c] for c in SENSE64])

# letter -> resolved bases (tools.c:14-16); U is T, '-', 'N' and '?' are any
_EQUATE = {"T": "T", "C": "C", "A": "A", "G": "G", "U": "T", "Y": "TC",
           "R": "AG", "M": "CA", "K": "TG", "S": "CG", "W": "TA", "H": "TCA",
           "B": "TCG", "V": "CAG", "D": "TAG", "-": "TCAG", "N": "TCAG",
           "?": "TCAG"}


def codon_str(i61):
    c = int(FROM61[i61])
    return BASES[c // 16] + BASES[(c // 4) % 4] + BASES[c % 4]


def resolve_letter(ch):
    try:
        return [BASES.index(b) for b in _EQUATE[ch.upper()]]
    except KeyError:
        raise ValueError(f"strange nucleotide {ch!r}")


def resolve_codon(triplet):
    """Sense codons (0..60) a possibly ambiguous triplet stands for."""
    sets = [resolve_letter(ch) for ch in triplet]
    out = []
    for b0 in sets[0]:
        for b1 in sets[1]:
            for b2 in sets[2]:
                s = FROM64[16 * b0 + 4 * b1 + b2]
                if s >= 0:
                    out.append(int(s))
    return sorted(out)


def position_bases(triplet):
    """Per codon position, the list of base indices the letter resolves to."""
    return [resolve_letter(ch) for ch in triplet]


def diff_tables():
    """(ndiff, pos, transition, synonymous) for every ordered pair of sense codons."""
    b = np.stack([FROM61 // 16, (FROM61 // 4) % 4, FROM61 % 4], axis=1)
    d = b[:, None, :] != b[None, :, :]
    ndiff = d.sum(-1)
    pos = np.argmax(d, axis=-1)
    fb = np.take_along_axis(np.broadcast_to(b[:, None, :], d.shape), pos[..., None], -1)[..., 0]
    tb = np.take_along_axis(np.broadcast_to(b[None, :, :], d.shape), pos[..., None], -1)[..., 0]
    s = fb + tb
    transition = (s == 1) | (s == 5)      # T<->C (0+1) or A<->G (2+3)
    syn = AA61[:, None] == AA61[None, :]
    return ndiff, pos, transition, syn
