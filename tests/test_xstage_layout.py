"""The LDS layout of the batched kernel's staged x chunks (XStage in
csrc/bhs_kernels.h), restated from its index formulas and checked for the two
properties the kernel relies on, for every pass width: the loads (instruction
i, lane l) cover every (item, chunk) of a pass exactly once, the pair reads
find each (item, chunk) where the loads put it, and any 16 consecutive lanes
reading one chunk of their items hit 16 different 16-byte bank groups."""
import pytest


def load_side(Cp, i, lane):
    """(item t, chunk c) that lane `lane` of load instruction i writes."""
    lc = Cp.bit_length() - 1
    lm = 4 - lc
    s = lane >> lc
    q = lane & (Cp - 1)
    w = s >> lm
    tb = 2 * (((w >> 1) << lm) + (s & ((1 << lm) - 1))) + (w & 1)
    t = tb + (i << (lm + 2))
    c = (q - 2 * i - (w >> 1)) & (Cp - 1)
    return t, c


def read_side(Cp, kk, lane, c):
    """(instruction, lane) block position that lane `lane` reads for its item kk
    at chunk c."""
    lc = Cp.bit_length() - 1
    lm = 4 - lc
    u = (kk >> 1) * 64 + lane
    r = u & ((1 << lm) - 1)
    v = u >> lm
    e = 2 * v + (kk & 1)
    s = r + ((e & 3) << lm)
    return e >> 2, (s << lc) + ((c + v) & (Cp - 1))


@pytest.mark.parametrize("Cp", [1, 2, 4, 8])
def test_loads_cover_each_item_chunk_once(Cp):
    seen = {}
    for i in range(4 * Cp):
        for lane in range(64):
            t, c = load_side(Cp, i, lane)
            assert 0 <= t < 256 and 0 <= c < Cp
            assert (t, c) not in seen
            seen[(t, c)] = (i, lane)
    assert len(seen) == 256 * Cp


@pytest.mark.parametrize("Cp", [1, 2, 4, 8])
def test_reads_find_what_the_loads_wrote(Cp):
    where = {}
    for i in range(4 * Cp):
        for lane in range(64):
            where[load_side(Cp, i, lane)] = (i, lane)
    for lane in range(64):
        for kk in range(4):
            t = (kk >> 1) * 128 + 2 * lane + (kk & 1)     # the tile order of item kk
            for c in range(Cp):
                assert read_side(Cp, kk, lane, c) == where[(t, c)]


@pytest.mark.parametrize("Cp", [1, 2, 4, 8])
def test_pair_reads_are_bank_conflict_free(Cp):
    for kk in range(4):
        for c in range(Cp):
            for base in range(0, 64, 16):
                groups = set()
                for lane in range(base, base + 16):
                    i, l = read_side(Cp, kk, lane, c)
                    groups.add(((1024 * i + 16 * l) // 16) % 16)
                assert len(groups) == 16, (Cp, kk, c, base)


def test_a_broken_rotation_is_caught():
    """Without the per-item chunk rotation the reads pile onto few bank groups."""
    Cp, kk, c = 8, 0, 3
    groups = set()
    for lane in range(16):
        i, l = read_side(Cp, kk, lane, c)
        l_flat = (l >> 3 << 3) + c                      # chunk slot without rotation
        groups.add(((1024 * i + 16 * l_flat) // 16) % 16)
    assert len(groups) < 16
