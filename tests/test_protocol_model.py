"""CPU model of the tile-status protocol (lookback, decoupled fallback, emulated
deadlock, epochs), run under random interleavings. No GPU.

The model mirrors the device code step by step (gpuprefixsums_amd/csrc/
device_common.hpp: status words, tickets, window lookback;
scan_u32.hip: csdldf_u32 and csdldf_pp_u32): every shared-memory access is
one step, and a seeded scheduler picks which resident tile (or persistent
workgroup) steps next. Tiles become resident in ticket order, at most
`resident` at a time; persistent workgroups are all resident and loop over
tickets. Several launches reuse one
status array, so stale words from earlier epochs are present. The checks:
each tile's exclusive prefix equals the true prefix mod 2^32, and across the
runs every fallback branch is taken (timeout, CAS win, CAS lost to an
aggregate, CAS lost to an inclusive word, fallback onto tile 0, stalled tile
released by a successor).
"""
import random
from collections import Counter

import pytest

MASK32 = 0xFFFFFFFF
AGG, INCL, CLAIM = 1, 2, 3
EPOCH_MASK = (1 << 30) - 1
CLAIM_PATIENCE = 64  # kClaimPatience in scan_u32.hip
REDUCTION_STEPS = 4  # a redundant reduction is long next to one status access


def make_status(value, flag, epoch):
    return (value & MASK32) | (flag << 32) | ((epoch & EPOCH_MASK) << 34)


def flag_of(w, epoch):
    """status_flag: AGG or INCL of this epoch, else 0 (stale or claimed)."""
    f = (w >> 32) & 3
    return f if (w >> 34) == (epoch & EPOCH_MASK) and f != CLAIM else 0


def claimed(w, epoch):
    return (w >> 34) == (epoch & EPOCH_MASK) and (w >> 32) & 3 == CLAIM


class Memory:
    def __init__(self, tiles):
        self.status = [0] * tiles
        self.ticket = 1 << 32

    def cas(self, i, expected, desired):
        found = self.status[i]
        if found == expected:
            self.status[i] = desired
        return found


def stall_then_publish(mem, t, agg, epoch, num_tiles, period, stats):
    """Generator: stall emulation, then the tile's aggregate publication."""
    # stall emulation: tiles t % period == 0 that have a successor wait until
    # their own word has been inserted by a successor
    if period and t % period == 0 and t + 1 < num_tiles:
        while True:
            w = mem.status[t]
            yield
            if flag_of(w, epoch):
                stats["stall_released"] += 1
                break
    mem.status[t] = make_status(agg[t], AGG if t else INCL, epoch)  # exchange
    yield


def lookback(mem, t, agg, epoch, spin_limit, stats):
    """Generator returning tile t's exclusive prefix (t > 0): window polls,
    fallback reductions and CAS insertions, one yield per access."""
    prefix, top, spins, claim_waits = 0, t - 1, 0, 0
    while True:
        # window read: lane l reads tile top - l (one step per word)
        words = []
        for l in range(min(64, top + 1)):
            words.append(mem.status[top - l])
            yield
        flags = [flag_of(w, epoch) for w in words]
        run = 0
        while run < len(flags) and flags[run]:
            run += 1
        first_incl = next((i for i, f in enumerate(flags) if f == INCL), len(flags))
        if first_incl < run:
            prefix += sum(w & MASK32 for w in words[:first_incl + 1])
            break
        if run:
            prefix += sum(w & MASK32 for w in words[:run])
            top -= run
            spins = 0
            continue
        spins += 1
        # patience grows with the distance to the blocking tile's nearest
        # successor (fallback_patience in scan_u32.hip)
        if spins < spin_limit * (1 + min(t - 1 - top, 7)):
            continue
        # claim_for_fallback: claim tile `top` before reducing it; a tile
        # already claimed is waited on for up to CLAIM_PATIENCE polls
        j, w0 = top, words[0]
        if claimed(w0, epoch):
            claim_waits += 1
            if claim_waits < CLAIM_PATIENCE:
                spins = 0
                stats["claim_wait"] += 1
                continue
            claim_waits, expect = 0, w0
            stats["claim_forced"] += 1
        else:
            claim = make_status(0, CLAIM, epoch)
            found = mem.cas(j, w0, claim)
            yield
            if found != w0:
                spins = 0
                stats["claim_lost"] += 1
                continue
            claim_waits, expect = 0, claim
        # fallback onto tile j: reduce its input, CAS over the claim
        stats["fallback"] += 1
        for _ in range(REDUCTION_STEPS):  # the redundant reduction of tile j's input
            yield
        mine = make_status(agg[j], AGG if j else INCL, epoch)
        found = mem.cas(j, expect, mine)
        yield
        stale = expect
        if found == stale:
            stats["cas_win"] += 1
            found, fl = mine, (AGG if j else INCL)
        else:
            fl = flag_of(found, epoch)
            assert fl, "a lost CAS must return a current-epoch word"
            stats["cas_lost_incl" if fl == INCL else "cas_lost_agg"] += 1
        if j == 0:
            stats["fallback_tile0"] += 1
        prefix += found & MASK32
        spins = 0
        if fl == INCL:
            break
        top = j - 1
    return prefix & MASK32


def resolve(mem, t, agg, epoch, spin_limit, stats, result):
    """Generator: tile t's lookback and inclusive publication."""
    if t == 0:
        result[t] = 0
        return
    prefix = yield from lookback(mem, t, agg, epoch, spin_limit, stats)
    assert flag_of(mem.status[t], epoch) == AGG, "the owner's word must be an aggregate before its inclusive one"
    mem.status[t] = make_status(prefix + agg[t], INCL, epoch)
    yield
    result[t] = prefix


def tile_program(mem, t, agg, epoch, num_tiles, period, spin_limit, stats, result):
    """Generator: one tile of the one-workgroup-per-tile decoupled-fallback
    kernel (csdldf_u32)."""
    yield from stall_then_publish(mem, t, agg, epoch, num_tiles, period, stats)
    yield from resolve(mem, t, agg, epoch, spin_limit, stats, result)


def workgroup_program(mem, tickets, agg, epoch, num_tiles, period, spin_limit, stats, result):
    """Generator: one workgroup of the persistent decoupled-fallback kernel
    (csdldf_pp_u32). It takes its next ticket right after publishing a
    tile's aggregate, then resolves that tile; a stalled tile therefore holds
    back the workgroup's next ticket."""
    t = tickets[0]
    tickets[0] += 1
    yield
    while t < num_tiles:
        yield from stall_then_publish(mem, t, agg, epoch, num_tiles, period, stats)
        nxt = tickets[0]
        tickets[0] += 1
        yield
        yield from resolve(mem, t, agg, epoch, spin_limit, stats, result)
        t = nxt


def is_stall_tile(t, num_tiles, period):
    return bool(period) and t % period == 0 and t + 1 < num_tiles


def workgroup_program_prepublish(mem, tickets, agg, epoch, num_tiles, period, spin_limit, stats, result):
    """Generator: the persistent decoupled-fallback kernel with stall-time
    prepublication. A workgroup whose tile t stalls first takes its next
    ticket and, unless that tile stalls too, publishes the next tile's
    aggregate before it waits, so nobody waits on the tile it holds; the
    stall is skipped when the next tile is t + 1 and the last tile (no other
    workgroup would be left to release it). Period 2 keeps the plain order:
    there the stalled workgroups could hold every successor between them."""
    t = tickets[0]
    tickets[0] += 1
    yield
    prepublished = False
    while t < num_tiles:
        if period >= 3 and is_stall_tile(t, num_tiles, period):
            assert not prepublished
            nxt = tickets[0]
            tickets[0] += 1
            yield
            pre_next = nxt < num_tiles and not is_stall_tile(nxt, num_tiles, period)
            if pre_next:
                mem.status[nxt] = make_status(agg[nxt], AGG, epoch)  # exchange; nxt > t >= 0
                stats["prepublished"] += 1
                yield
            if not (nxt == t + 1 and nxt + 1 == num_tiles):
                while True:
                    w = mem.status[t]
                    yield
                    if flag_of(w, epoch):
                        stats["stall_released"] += 1
                        break
            else:
                stats["stall_skipped"] += 1
            mem.status[t] = make_status(agg[t], AGG if t else INCL, epoch)  # exchange
            yield
            yield from resolve(mem, t, agg, epoch, spin_limit, stats, result)
            prepublished = pre_next
            t = nxt
            continue
        if not prepublished:
            yield from stall_then_publish(mem, t, agg, epoch, num_tiles, period, stats)
        nxt = tickets[0]
        tickets[0] += 1
        yield
        yield from resolve(mem, t, agg, epoch, spin_limit, stats, result)
        prepublished = False
        t = nxt


def run_launch(mem, agg, epoch, period, spin_limit, resident, rng, stats):
    num_tiles = len(agg)
    result = [None] * num_tiles
    progs, next_tile, active = {}, 0, []
    steps = 0
    while next_tile < num_tiles or active:
        while len(active) < resident and next_tile < num_tiles:  # dispatch in ticket order
            progs[next_tile] = tile_program(mem, next_tile, agg, epoch, num_tiles, period, spin_limit, stats, result)
            active.append(next_tile)
            next_tile += 1
        t = rng.choice(active)
        try:
            next(progs[t])
        except StopIteration:
            active.remove(t)
        steps += 1
        assert steps < 5_000_000, "protocol did not terminate"
    return result


def run_persistent(mem, agg, epoch, period, spin_limit, grid, rng, stats, program=None):
    """All `grid` workgroups are resident from the start (the persistent
    grid is sized to what is resident) and share one ticket counter."""
    num_tiles = len(agg)
    result = [None] * num_tiles
    tickets = [0]
    program = program or workgroup_program
    progs = {g: program(mem, tickets, agg, epoch, num_tiles, period, spin_limit, stats, result)
             for g in range(min(grid, num_tiles))}
    active = list(progs)
    steps = 0
    while active:
        g = rng.choice(active)
        try:
            next(progs[g])
        except StopIteration:
            active.remove(g)
        steps += 1
        assert steps < 5_000_000, "protocol did not terminate"
    assert tickets[0] == num_tiles + min(grid, num_tiles)  # every workgroup ends on one ticket past the tiles
    return result


def run_split(mem, agg, epoch, period, spin_limit, resident, rng, stats):
    """The split schedule (GPS_SCHEDULE_SPLIT, the D3D12 deadlock emulation):
    a first launch of every tile with index % period != 0, in order, then a
    second launch of the deferred tiles under the same epoch. Returns the
    prefixes and, as of the end of the first launch, the deferred tiles whose
    word a fallback had inserted (all of them had not started)."""
    num_tiles = len(agg)
    result = [None] * num_tiles
    first = [t for t in range(num_tiles) if t % period]
    second = [t for t in range(num_tiles) if t % period == 0]
    inserted_before_start = []
    for launch in (first, second):
        progs, active, i, steps = {}, [], 0, 0
        while i < len(launch) or active:
            while len(active) < resident and i < len(launch):  # dispatch in ticket order
                t = launch[i]
                progs[t] = tile_program(mem, t, agg, epoch, num_tiles, 0, spin_limit, stats, result)
                active.append(t)
                i += 1
            t = rng.choice(active)
            try:
                next(progs[t])
            except StopIteration:
                active.remove(t)
            steps += 1
            assert steps < 5_000_000, "protocol did not terminate"
        if launch is first:
            inserted_before_start = [t for t in second if flag_of(mem.status[t], epoch)]
    return result, inserted_before_start


@pytest.mark.parametrize("seed", range(12))
def test_split_schedule_under_random_interleavings(seed):
    # Deferred tiles never run during the first launch, so the fallback must
    # reduce and insert every one of them that has a successor; the second
    # launch then completes them. Every prefix is exact.
    rng = random.Random(700 + seed)
    num_tiles = rng.choice([2, 3, 9, 40, 130])
    mem = Memory(200)
    for launch in range(3):
        epoch = launch + 1
        agg = [rng.getrandbits(32) for _ in range(num_tiles)]
        period = rng.choice([2, 3, 7, 8])
        stats = Counter()
        res, inserted = run_split(mem, agg, epoch, period, spin_limit=rng.choice([1, 2, 4]),
                                  resident=rng.choice([2, 5, 17]), rng=rng, stats=stats)
        acc = 0
        for t in range(num_tiles):
            assert res[t] == acc, (seed, launch, t)
            acc = (acc + agg[t]) & MASK32
        deferred_with_successor = [t for t in range(0, num_tiles, period) if t + 1 < num_tiles]
        assert inserted == deferred_with_successor
        assert stats["cas_win"] >= len(deferred_with_successor)
        if 0 in deferred_with_successor:
            assert stats["fallback_tile0"] > 0
        assert all(flag_of(mem.status[t], epoch) == INCL for t in range(num_tiles))


def test_split_tile_mapping_matches_the_device_formula():
    # split_tile / split_tile_count in device_common.hpp: ticket q of the
    # first launch is the q-th tile with index % P != 0, of the second q * P.
    for num_tiles in (2, 3, 7, 8, 9, 64, 1000):
        for P in (2, 3, 7, 8):
            deferred = (num_tiles + P - 1) // P
            first = [q + q // (P - 1) + 1 for q in range(num_tiles - deferred)]
            second = [q * P for q in range(deferred)]
            assert first == [t for t in range(num_tiles) if t % P]
            assert second == [t for t in range(num_tiles) if t % P == 0]


def test_prev_epoch_inverts_next_epoch():
    def prev_epoch(e):  # device_common.hpp, on masked epochs
        p = (e - 1) & EPOCH_MASK
        return EPOCH_MASK if p == 0 else p
    for raw in (1, 2, 5, EPOCH_MASK - 1, EPOCH_MASK, EPOCH_MASK + 2, (7 << 30) + 3):
        nxt = next_epoch(raw)
        assert prev_epoch(nxt & EPOCH_MASK) == raw & EPOCH_MASK


@pytest.mark.parametrize("seed", range(12))
def test_protocol_under_random_interleavings(seed):
    rng = random.Random(seed)
    num_tiles = rng.choice([1, 2, 9, 40, 130])
    mem = Memory(200)
    stats = Counter()
    for launch in range(4):  # reuse the status array: stale words from older epochs
        epoch = launch + 1
        agg = [rng.getrandbits(32) for _ in range(num_tiles)]
        period = rng.choice([0, 2, 7, 8])
        res = run_launch(mem, agg, epoch, period, spin_limit=rng.choice([1, 2, 4]),
                         resident=rng.choice([2, 5, 17]), rng=rng, stats=stats)
        acc = 0
        for t in range(num_tiles):
            assert res[t] == acc, (seed, launch, t)
            acc = (acc + agg[t]) & MASK32
        # after the launch every word is inclusive for this epoch
        assert all(flag_of(mem.status[t], epoch) == INCL for t in range(num_tiles))


def test_every_fallback_branch_is_exercised():
    stats = Counter()
    for seed in range(40):
        rng = random.Random(1000 + seed)
        mem = Memory(64)
        agg = [rng.getrandbits(32) for _ in range(48)]
        res = run_launch(mem, agg, 1 + seed % 3, period=rng.choice([2, 7, 8]), spin_limit=1,
                         resident=rng.choice([3, 9]), rng=rng, stats=stats)
        acc = 0
        for t, r in enumerate(res):
            assert r == acc
            acc = (acc + agg[t]) & MASK32
    for branch in ("fallback", "cas_win", "fallback_tile0", "stall_released", "claim_lost", "claim_wait"):
        assert stats[branch] > 0, (branch, dict(stats))
    # waiters behind a claimed tile wait for its insertion instead of reducing
    # it again: every reduction inserts
    assert stats["fallback"] == stats["cas_win"] + stats["cas_lost_agg"] + stats["cas_lost_incl"]
    assert stats["fallback"] <= 1.05 * stats["cas_win"]


def test_forced_fallback_on_a_long_claim(monkeypatch):
    # With a tiny claim patience, waiters stop waiting on a claimed tile and
    # reduce it themselves; the results stay exact.
    monkeypatch.setitem(globals(), "CLAIM_PATIENCE", 2)
    monkeypatch.setitem(globals(), "REDUCTION_STEPS", 40)
    stats = Counter()
    for seed in range(20):
        rng = random.Random(3000 + seed)
        mem = Memory(64)
        agg = [rng.getrandbits(32) for _ in range(48)]
        res = run_launch(mem, agg, 1 + seed % 3, period=rng.choice([2, 7]), spin_limit=1,
                         resident=rng.choice([5, 9]), rng=rng, stats=stats)
        acc = 0
        for t, r in enumerate(res):
            assert r == acc
            acc = (acc + agg[t]) & MASK32
    assert stats["claim_forced"] > 0, dict(stats)
    # a forced reduction's insertion loses to the claimer's (or the owner's) word
    assert stats["cas_lost_agg"] + stats["cas_lost_incl"] > 0, dict(stats)


def test_stale_epoch_words_are_never_consumed():
    # A status array full of inclusive words from an older epoch must look
    # "not ready" to the new launch: results stay exact.
    rng = random.Random(7)
    mem = Memory(50)
    mem.status = [make_status(rng.getrandbits(32), INCL, 5) for _ in range(50)]
    agg = [rng.getrandbits(32) for _ in range(50)]
    stats = Counter()
    res = run_launch(mem, agg, 6, period=0, spin_limit=3, resident=4, rng=rng, stats=stats)
    acc = 0
    for t, r in enumerate(res):
        assert r == acc
        acc = (acc + agg[t]) & MASK32


@pytest.mark.parametrize("seed", range(12))
def test_persistent_schedule_under_random_interleavings(seed):
    # The persistent kernel's stalls hold back a workgroup's next ticket; the
    # launch must still finish (no deadlock) with every prefix exact.
    rng = random.Random(500 + seed)
    num_tiles = rng.choice([1, 2, 9, 40, 130])
    mem = Memory(200)
    stats = Counter()
    for launch in range(4):
        epoch = launch + 1
        agg = [rng.getrandbits(32) for _ in range(num_tiles)]
        res = run_persistent(mem, agg, epoch, period=rng.choice([0, 2, 3, 7, 8]), spin_limit=rng.choice([1, 2, 4]),
                             grid=rng.choice([2, 5, 17]), rng=rng, stats=stats)
        acc = 0
        for t in range(num_tiles):
            assert res[t] == acc, (seed, launch, t)
            acc = (acc + agg[t]) & MASK32
        assert all(flag_of(mem.status[t], epoch) == INCL for t in range(num_tiles))


@pytest.mark.parametrize("seed", range(16))
def test_prepublishing_persistent_schedule_under_random_interleavings(seed):
    # Stall-time prepublication: the launch must finish with every prefix
    # exact, at every period (2 keeps the plain order), grid and tile count,
    # including the tile just before the last one stalling.
    rng = random.Random(900 + seed)
    num_tiles = rng.choice([1, 2, 3, 4, 9, 40, 130])
    mem = Memory(200)
    stats = Counter()
    for launch in range(4):
        epoch = launch + 1
        agg = [rng.getrandbits(32) for _ in range(num_tiles)]
        res = run_persistent(mem, agg, epoch, period=rng.choice([0, 2, 3, 4, 7, 8]), spin_limit=rng.choice([1, 2, 4]),
                             grid=rng.choice([2, 3, 5, 17]), rng=rng, stats=stats,
                             program=workgroup_program_prepublish)
        acc = 0
        for t in range(num_tiles):
            assert res[t] == acc, (seed, launch, t)
            acc = (acc + agg[t]) & MASK32
        assert all(flag_of(mem.status[t], epoch) == INCL for t in range(num_tiles))


def test_prepublishing_takes_its_branches_and_cuts_insertions():
    """Across many launches: tiles are prepublished, the last-but-one stall
    is skipped when its workgroup holds the last tile, and fewer tiles that
    never stalled are reduced by a waiter than in the plain order."""
    plain, pre = Counter(), Counter()
    for seed in range(40):
        for program, stats in ((workgroup_program, plain), (workgroup_program_prepublish, pre)):
            rng = random.Random(3000 + seed)
            num_tiles = rng.choice([5, 9, 24, 48])  # 5 and 9: the tile before the last stalls at period 3 or 7
            agg = [rng.getrandbits(32) for _ in range(num_tiles)]
            res = run_persistent(Memory(64), agg, 1, period=rng.choice([3, 7]), spin_limit=1,
                                 grid=rng.choice([2, 3, 9]), rng=rng, stats=stats, program=program)
            acc = 0
            for t, r in enumerate(res):
                assert r == acc
                acc = (acc + agg[t]) & MASK32
    assert pre["prepublished"] > 0 and pre["stall_skipped"] > 0, dict(pre)
    assert pre["cas_win"] + pre["cas_lost_agg"] <= plain["cas_win"] + plain["cas_lost_agg"], (dict(pre), dict(plain))


def test_persistent_schedule_exercises_the_fallback_branches():
    stats = Counter()
    for seed in range(30):
        rng = random.Random(2000 + seed)
        mem = Memory(64)
        agg = [rng.getrandbits(32) for _ in range(48)]
        res = run_persistent(mem, agg, 1 + seed % 3, period=rng.choice([2, 3, 7, 8]), spin_limit=1,
                             grid=rng.choice([2, 3, 9]), rng=rng, stats=stats)
        acc = 0
        for t, r in enumerate(res):
            assert r == acc
            acc = (acc + agg[t]) & MASK32
    for branch in ("fallback", "cas_win", "cas_lost_agg", "fallback_tile0", "stall_released"):
        assert stats[branch] > 0, (branch, dict(stats))


# ---------------------------------------------------------------------------
# Ticket-word accounting (device_common.hpp: take_ticket, drain_tickets).
# The word holds (epoch << 32) | tickets_taken; a persistent launch takes
# num_tiles + grid tickets (each workgroup ends on one past the last tile) and
# the holder of the last re-arms the word to (next epoch, 0). A workgroup that
# gives up early (a watchdog) drains tickets until it holds one past the last
# tile, so the count still reaches the total.
# ---------------------------------------------------------------------------
def next_epoch(e):
    n = e + 1
    if (n & EPOCH_MASK) == 0:
        n += 1
    return n


class TicketWord:
    def __init__(self, epoch):
        self.v = epoch << 32

    def fetch_add(self):
        old = self.v
        self.v += 1
        return old

    def rearm_if_last(self, old, total):
        if (old & MASK32) == total - 1:
            self.v = next_epoch(old >> 32) << 32


def persistent_tickets(word, num_tiles, total, give_up_after, drain):
    """Generator for one workgroup: takes tickets until one past the last tile;
    with give_up_after = k it stops after holding its k-th ticket, draining
    (drain_tickets) or simply returning."""
    taken = 0
    while True:
        old = word.fetch_add()
        word.rearm_if_last(old, total)
        taken += 1
        yield
        if (old & MASK32) >= num_tiles:
            return
        if give_up_after is not None and taken >= give_up_after:
            if drain:
                t = old
                while (t & MASK32) < num_tiles:
                    t = word.fetch_add()
                    word.rearm_if_last(t, total)
                    yield
            return


def run_ticket_launch(word, num_tiles, grid, give_up, rng, drain=True):
    total = num_tiles + grid
    progs = [persistent_tickets(word, num_tiles, total, give_up.get(g), drain) for g in range(grid)]
    while progs:
        p = rng.choice(progs)
        try:
            next(p)
        except StopIteration:
            progs.remove(p)


@pytest.mark.parametrize("seed", range(20))
def test_ticket_word_rearms_after_early_exits(seed):
    rng = random.Random(9000 + seed)
    epoch = rng.choice([1, 7, EPOCH_MASK - 1, EPOCH_MASK])
    word = TicketWord(epoch)
    for launch in range(5):
        num_tiles, grid = rng.choice([1, 3, 40, 200]), rng.choice([1, 2, 8, 33])
        give_up = {g: rng.randint(1, 4) for g in range(grid) if rng.random() < 0.3}
        run_ticket_launch(word, num_tiles, grid, give_up, rng)
        epoch = next_epoch(epoch)
        assert word.v == epoch << 32, (seed, launch, hex(word.v))


def test_early_exit_without_draining_leaves_the_word_armed_wrong():
    # The failure the drain prevents: a workgroup that returns while holding a
    # tile leaves the count short, so the word is never re-armed.
    rng = random.Random(1)
    word = TicketWord(3)
    run_ticket_launch(word, 40, 4, {0: 1}, rng, drain=False)
    assert word.v >> 32 == 3 and (word.v & MASK32) < 44
