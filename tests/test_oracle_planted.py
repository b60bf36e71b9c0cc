"""The known-answer tests notice a flipped rule.  Each planted variant of the
oracle (fot_oracle.c compiled with ORACLE_PLANT=k, `make -C oracle planted`)
changes exactly one rule of the restatement; the known answers written for
that rule must fail on it, and every other known answer must still pass, so
each answer pins its own rule.  CPU only; the variants are never used as a
checker."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from kats import (HPLAN_KATS, MINE_KATS, kat_global_colour_least_used,
                  kat_hplan_block_colour_least_used, kat_hplan_block_split,
                  kat_hplan_thread_order, kat_mine2, kat_mine4, kat_mine8)

PLANTED_DIR = os.path.join(O.HERE, "lib", "planted")

# ORACLE_PLANT value -> the known answers that must catch it (the enum in
# fot_oracle.c names each flipped rule)
CAUGHT_BY = {
    1: [kat_hplan_block_colour_least_used],  # block colour by first fit
    2: [kat_hplan_block_split],              # no count restart at a part change
    3: [kat_hplan_thread_order, kat_hplan_block_split],  # positions descending in a
                                                         # thread colour
    4: [kat_hplan_thread_order, kat_hplan_block_colour_least_used,
        kat_hplan_block_split],              # cached points descending
    5: [kat_mine2],                          # mine<2> without its factor 2
    6: [kat_mine2],                          # mine<2> out_L from L
    7: [kat_mine4, kat_mine8],               # mine<4/8> wrong ring neighbour
    8: [kat_global_colour_least_used],       # global colouring by first fit
}
ALL = [kat_global_colour_least_used] + HPLAN_KATS + MINE_KATS


def _impl(lib):
    def hplan(maps, npts, bs, part):
        return O.OraclePlan(maps, npts, bs, part, library=lib).colours()

    def run(kernel, maps, npts, pd, cd, iters, soa):
        o = O.OracleProblem(kernel, maps, npts, [pd], cd, soa=soa, library=lib)
        o.loop_serial(iters)
        return o.target

    def colour(maps0, npts):
        return O.colour_cells(maps0, npts, library=lib)
    return {kat_global_colour_least_used: colour,
            **{k: hplan for k in HPLAN_KATS}, **{k: run for k in MINE_KATS}}


@pytest.fixture(scope="module")
def planted():
    subprocess.run(["make", "-C", O.HERE, "planted"], check=True, stdout=subprocess.DEVNULL)
    return {k: O.load(os.path.join(PLANTED_DIR, f"liboracle_plant{k}.so")) for k in CAUGHT_BY}


@pytest.mark.parametrize("kat", ALL, ids=[k.__name__ for k in ALL])
def test_known_answer_holds_for_the_oracle(kat):
    kat(_impl(O.lib)[kat])


@pytest.mark.parametrize("plant", sorted(CAUGHT_BY))
def test_planted_mistake_is_caught(planted, plant):
    impl = _impl(planted[plant])
    for kat in ALL:
        if kat in CAUGHT_BY[plant]:
            with pytest.raises(AssertionError):
                kat(impl[kat])
        else:
            kat(impl[kat])  # the other rules are untouched
