"""CPU emulation of the plane bookkeeping of the multi-plane tile steps in csrc/grid.hip.

A block of a tiled pass marches over planes [i0, i1) of one level (nx planes) in steps of PL
planes and keeps staged planes in a ring of LDS slots.  The emulation follows each kernel's
loads, slot stores and per-step computations with plane labels instead of values, and checks
that every plane of the chunk is computed exactly once, each time from the slots holding the
planes its stencil reads (or zero outside the level box), for the chunk sizes and step widths
the kernels are built with and for chunks that end on a partial step.
"""
import itertools

import pytest

ZERO = "zero"


def plane(i, nx, hi):
    """What a guarded load of plane i yields: the plane inside the box (and up to hi), else 0."""
    return i if 0 <= i < nx and i <= hi else ZERO


def emulate_v(i0, i1, nx, PL, NS=None):
    """sv_tile: V on planes i..i+PL-1 of a step reads x planes ii-1, ii, ii+1 from a ring of
    PL + 2 slots; the step's loads (issued a step ahead) are x planes i+1..i+PL."""
    NS = NS or PL + 2
    slots = [None] * NS
    slots[(i0 - 1 + NS) % NS] = plane(i0 - 1, nx, i1)
    slots[i0 % NS] = plane(i0, nx, i1)
    pending = [plane(i0 + l + 1, nx, i1) for l in range(PL)]
    done = []
    for i in range(i0, i1, PL):
        for l in range(PL):
            slots[(i + l + 1) % NS] = pending[l]
        if i + PL < i1:
            pending = [plane(i + PL + l + 1, nx, i1) for l in range(PL)]
        for l in range(PL):
            ii = i + l
            if ii < i1:
                got = (slots[(ii - 1 + NS) % NS], slots[ii % NS], slots[(ii + 1) % NS])
                want = tuple(plane(q, nx, i1) for q in (ii - 1, ii, ii + 1))
                assert got == want, (i0, i1, nx, PL, ii, got, want)
                done.append(ii)
    return done


def emulate_g1(i0, i1, nx, PL):
    """sg1_tile: G1 on plane ii reads the x-edges of plane ii (slot l of the step) and the y/z
    edges of planes ii and ii+1 from a ring of PL + 1 slots; a step's loads are x-edges of
    planes i..i+PL-1 and y/z-edges of planes i+1..i+PL (x-edge planes < i1, y/z planes <= i1)."""
    NS = PL + 1
    yz = [None] * NS
    yz[i0 % NS] = i0 if i0 <= nx else ZERO
    def fetch(i):
        return ([i + l if i + l < i1 else ZERO for l in range(PL)],
                [i + l + 1 if i + l + 1 <= i1 else ZERO for l in range(PL)])
    pend = fetch(i0)
    done = []
    for i in range(i0, i1, PL):
        sx = list(pend[0])
        for l in range(PL):
            yz[(i + l + 1) % NS] = pend[1][l]
        if i + PL < i1:
            pend = fetch(i + PL)
        for l in range(PL):
            ii = i + l
            if ii < i1:
                assert sx[l] == ii, (i0, i1, PL, ii, sx)
                assert (yz[ii % NS], yz[(ii + 1) % NS]) == (ii, ii + 1), (i0, i1, PL, ii, yz)
                done.append(ii)
    return done


def emulate_g3(i0, i1, nx, PL):
    """sg3_tile: the edges of plane ii read y planes ii and ii-1 from a ring of PL + 1 slots;
    a step's loads are y planes i..i+PL-1 (planes < i1) and plane i0-1 is staged first."""
    NS = PL + 1
    slots = [None] * NS
    slots[(i0 - 1 + NS) % NS] = plane(i0 - 1, nx, i1)
    pend = [plane(i0 + l, nx, i1) if i0 + l < i1 else None for l in range(PL)]
    done = []
    for i in range(i0, i1, PL):
        for l in range(PL):
            if i + l < i1:
                slots[(i + l + NS) % NS] = pend[l]
        if i + PL < i1:
            pend = [plane(i + PL + l, nx, i1) if i + PL + l < i1 else None for l in range(PL)]
        for l in range(PL):
            ii = i + l
            if ii < i1:
                got = (slots[ii % NS], slots[(ii - 1 + NS) % NS])
                assert got == (plane(ii, nx, i1), plane(ii - 1, nx, i1)), (i0, i1, PL, ii, got)
                done.append(ii)
    return done


@pytest.mark.parametrize("kind,PL", [("v", 1), ("v", 2), ("v", 4), ("g1", 1), ("g1", 2), ("g1", 4),
                                     ("g3", 1), ("g3", 2), ("g3", 4)])
def test_every_plane_once_from_the_right_slots(kind, PL):
    emulate = {"v": emulate_v, "g1": emulate_g1, "g3": emulate_g3}[kind]
    for nx, ci in itertools.product((1, 2, 3, 5, 7, 8, 16, 19, 48), (8, 16, 32)):
        for i0 in range(0, nx, ci):
            i1 = min(i0 + ci, nx)
            assert emulate(i0, i1, nx, PL) == list(range(i0, i1))


def test_a_wrong_ring_size_is_caught():
    """A V ring one slot short (a plausible slip) hands a plane the wrong neighbour."""
    with pytest.raises(AssertionError):
        emulate_v(0, 16, 16, 4, NS=5)
