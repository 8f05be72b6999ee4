"""The wavefront render's workspace layout and its two-stream schedule, on
the CPU.

1. Layout (through rt_wavefront_state_regions, i.e. the library's own
   wf_layout / wf_state): for 2..4 sample slots and several image and batch
   sizes, every buffer a slot's state addresses lies inside the workspace,
   starts on a 256-byte boundary, and is disjoint from every other buffer;
   shared buffers are the same for every slot; a slot's own buffers and its
   shadow count are disjoint from every other slot's.
2. Schedule: a model of the launches and events rt_launch_wavefront
   (cuda-rt_amd/csrc/rt_trace.hip) enqueues -- which passes run on the
   caller's stream and which on the auxiliary one, which buffers each pass
   reads and writes (the WfState fields its kernel uses), and the event
   record / wait edges. Every pair of passes on different streams that touch
   a common non-atomic buffer, one of them writing, must be ordered by the
   happens-before relation of stream order plus event edges, and every
   slot's shadow-sort histogram must be cleared before its first sort. The
   model passes for 2, 3 and 4 slots, and it catches the aliasing of two
   in-flight samples onto one slot's buffers (a schedule with three samples
   in flight over two slots' buffers), the kind of mistake that made an
   earlier three-slot variant fault.
"""
import ctypes
import itertools

import pytest

import rtamd


class Region(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("offset", ctypes.c_uint64), ("bytes", ctypes.c_uint64),
                ("shared", ctypes.c_int32)]


def regions(npix, ntiles, nscenes, nslots, slot):
    L = rtamd.lib()
    f = L.rt_wavefront_state_regions
    f.argtypes = [ctypes.c_uint64] * 3 + [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Region), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_uint64)]
    f.restype = ctypes.c_int32
    out = (Region * 64)()
    total = ctypes.c_uint64()
    n = f(npix, ntiles, nscenes, nslots, slot, out, 64, ctypes.byref(total))
    assert 0 < n <= 64
    return {r.name.decode(): (int(r.offset), int(r.bytes), bool(r.shared)) for r in out[:n]}, int(total.value)


SIZES = [(1, 1, 1), (64 * 48, 12, 1), (1920 * 1080, 120 * 68, 1), (5_000_000, 20_000, 64), (300_000, 1200, 300)]


@pytest.mark.parametrize("npix,ntiles,nscenes", SIZES)
@pytest.mark.parametrize("nslots", [2, 3, 4])
def test_slot_buffers_in_bounds_aligned_and_disjoint(npix, ntiles, nscenes, nslots):
    per = []
    for slot in range(nslots):
        reg, total = regions(npix, ntiles, nscenes, nslots, slot)
        per.append(reg)
        for name, (off, size, _) in reg.items():
            assert 0 <= off and off + size <= total, (slot, name)
            if name != "scount":
                assert off % 256 == 0, (slot, name)
    shared = {k: v for k, v in per[0].items() if v[2]}
    for reg in per[1:]:
        assert {k: v for k, v in reg.items() if v[2]} == shared
    # every buffer (shared, and each slot's own) against every other
    spans = [("shared", k, v[0], v[1]) for k, v in shared.items() if k != "scount"]
    for slot, reg in enumerate(per):
        spans += [(slot, k, v[0], v[1]) for k, v in reg.items() if not v[2] and k != "scount"]
    spans.sort(key=lambda s: s[2])
    for a, b in zip(spans, spans[1:]):
        assert a[2] + a[3] <= b[2], (a, b)
    # shadow counts: inside counts, one word per slot, clear of the queue
    # counts [0], [1] and the batch overflow count [8]
    c_off, c_size, _ = shared["counts"]
    words = [(reg["scount"][0] - c_off) // 4 for reg in per]
    assert len(set(words)) == nslots and all(2 <= w < 8 for w in words)
    assert all(c_off <= reg["scount"][0] < c_off + c_size for reg in per)


def test_render_layout_is_two_slots():
    reg2, total2 = regions(1920 * 1080, 120 * 68, 1, 2, 0)
    reg3, total3 = regions(1920 * 1080, 120 * 68, 1, 3, 0)
    assert total3 > total2
    # the library refuses slot counts it does not lay out
    L = rtamd.lib()
    f = L.rt_wavefront_state_regions
    assert f(100, 1, 1, 7, 0, None, 0, None) == -1
    assert f(100, 1, 1, 2, 2, None, 0, None) == -1


# ---------------------------------------------------------------- schedule model
ATOMIC = {"rep", "ovf"}


def schedule(spp, depth, nslots, slot_buffers=None, pair=None):
    """The passes rt_launch_wavefront enqueues, as (stream, name, reads,
    writes) plus ('record', event) / ('wait', event) markers, in enqueue
    order. slot_buffers(it) names the buffer set sample `it` uses (default
    it % nslots, as the code does); the wait before sample it's first
    shading pass is on the shadow finish of sample it - nslots, whose
    shadow count and any-hit fetch counters are then cleared and whose path
    sums that shading pass adds to the image; the last nslots samples are
    added after the auxiliary stream is done. With pair (the code's choice
    for two slots), the last bounce's any-hit pass of an even sample waits
    for the next sample's and runs in one launch with it, followed by both
    samples' shadow finishes; both flushes are recorded after them."""
    slot_of = slot_buffers or (lambda it: it % nslots)
    if pair is None:
        pair = nslots == 2
    ops = []

    def op(stream, name, reads=(), writes=()):
        ops.append(("op", stream, name, set(reads), set(writes)))

    def own(k, it):
        return f"{k}[{slot_of(it)}]"

    ops.append(("record", "s", "start"))
    ops.append(("wait", "a", "start"))
    op("a", "shade_fill", ["scenes"], ["shade_table"])
    ops.append(("record", "a", "filled"))
    op("a", "zero color", [], ["color"])
    op("s", "init", [], ["counts01", "rep", "shist[0]", "rng"] + [f"scount[{j}]" for j in range(nslots)])
    op("s", "reset fetch0", [], ["fetch0"])
    for j in range(nslots):
        op("s", f"reset fetch1[{j}]", [], [f"fetch1[{j}]"])
        if j > 0:
            op("s", f"reset shist[{j}]", [], [f"shist[{j}]"])
    flushed = {}
    for it in range(spp):
        op("s", f"camera {it}", ["rng"], ["rng", "ray0", "pix0", "rscene0", "counts01"])
        cur = 0
        for b in range(depth + 1):
            ray, nxt = f"ray{cur}", f"ray{cur ^ 1}"
            op("s", f"closest {it}.{b}", [ray, "counts01", "views"], ["hrec", "fetch0", "rep", "ovf"])
            if it == 0 and b == 0:
                ops.append(("wait", "s", "filled"))
            flush = b == 0 and it >= nslots
            if flush:
                ops.append(("wait", "s", ("flushed", it - nslots)))
                op("s", f"reset slot {it}", [], [own("scount", it), own("fetch1", it)])
            op("s", f"shade {it}.{b}", ["hrec", ray, f"pix{cur}", f"rscene{cur}", "thr", "rng", "shade_table",
                                         own("scount", it), own("acc", it)] + (["color"] if flush else []),
               ["rng", "thr", own("acc", it), nxt, f"pix{cur ^ 1}", f"rscene{cur ^ 1}", "counts01", own("shq", it),
                own("spix", it), own("sscene", it), own("scount", it), own("skey", it), "rep", f"rec{it}"]
               + (["color"] if flush else []))
            op("s", f"reset queue {it}.{b}", [], ["counts01", "fetch0"])
            sb = "s"
            if b == depth:
                ops.append(("record", "s", ("shaded", it)))
                ops.append(("wait", "a", ("shaded", it)))
                sb = "a"
            op(sb, f"sort {it}.{b}", [own("shq", it), own("scount", it), own("skey", it)],
               [own("skey", it), own("shist", it), own("sbase", it), own("sperm", it), own("srank", it)])
            if b == depth and pair and it % 2 == 0 and it + 1 < spp:
                cur ^= 1  # deferred: traced with sample it + 1's shadow rays
                continue
            group = [it - 1, it] if b == depth and pair and it % 2 == 1 else [it]
            # a paired launch fetches from both samples' queues and uses the
            # first one's fetch counters
            op(sb, f"any {it}.{b}", sum(([own("shq", k), own("sperm", k), own("scount", k), own("sscene", k)]
                                         for k in group), ["views"]),
               [own("srec", k) for k in group] + [own("fetch1", group[0]), "rep", "ovf"])
            for k in group:
                op(sb, f"shadow finish {k}.{b}", [own("srec", k), own("srank", k), own("shq", k), own("spix", k),
                                                 own("sscene", k), own("scount", k)],
                   [own("acc", k), "rep", f"rec{k}"])
            if b < depth:
                op(sb, f"reset shadow {it}.{b}", [], [own("scount", it), own("fetch1", it)])
            cur ^= 1
        if pair and it % 2 == 0 and it + 1 < spp:
            continue
        if pair and it % 2 == 1:
            ops.append(("record", "a", ("flushed", it - 1)))
        ops.append(("record", "a", ("flushed", it)))
    ops.append(("record", "a", "done"))
    ops.append(("wait", "s", "done"))
    op("s", "flush last", [own("acc", it) for it in range(max(0, spp - nslots), spp)] + ["color"], ["color"])
    op("s", "resolve", ["color"], ["out"])
    return ops


def hazards(ops):
    """Unordered conflicting pass pairs, and sorts of a slot whose histogram
    no earlier pass cleared, under stream order + event edges (a wait sees
    the latest record of its event enqueued before it)."""
    passes, last_on, last_record = [], {}, {}
    preds = []  # happens-before predecessors (indices into passes) of each pass
    pending_waits = {"s": set(), "a": set()}
    for o in ops:
        if o[0] == "record":
            last_record[o[2]] = last_on.get(o[1])
        elif o[0] == "wait":
            r = last_record.get(o[2])
            if r is not None:
                pending_waits[o[1]].add(r)
        else:
            _, stream, name, reads, writes = o
            p = set(pending_waits[stream])
            pending_waits[stream] = set()
            if stream in last_on:
                p.add(last_on[stream])
            idx = len(passes)
            passes.append((stream, name, reads, writes))
            preds.append(p)
            last_on[stream] = idx
    # transitive closure (passes are in enqueue order, so predecessors precede)
    before = []
    for i, p in enumerate(preds):
        s = set()
        for j in p:
            s |= before[j]
            s.add(j)
        before.append(s)
    bad = []
    for i, j in itertools.combinations(range(len(passes)), 2):
        si, _, ri, wi = passes[i]
        sj, _, rj, wj = passes[j]
        if si == sj:
            continue
        common = ((wi & (rj | wj)) | (wj & ri)) - ATOMIC
        if common and i not in before[j] and j not in before[i]:
            bad.append((passes[i][1], passes[j][1], sorted(common)))
    unclear = []
    for i, (_, name, _, writes) in enumerate(passes):
        if name.startswith("sort"):
            hist = [w for w in writes if w.startswith("shist")][0]
            if not any(hist in passes[j][3] and not passes[j][1].startswith("sort") for j in before[i]):
                unclear.append(name)
    return bad, unclear


@pytest.mark.parametrize("pair", [False, True])
@pytest.mark.parametrize("nslots", [2, 3, 4])
@pytest.mark.parametrize("spp,depth", [(1, 0), (5, 0), (6, 0), (7, 2)])
def test_schedule_orders_every_conflict(nslots, spp, depth, pair):
    bad, unclear = hazards(schedule(spp, depth, nslots, pair=pair))
    assert not bad, bad[:5]
    assert not unclear, unclear


def test_model_catches_a_pair_flushed_before_its_finish():
    # sample 0's flush event recorded when its own passes are queued, not
    # after the paired any-hit launch and shadow finish: sample 2's shading
    # pass could add (and restart) slot 0's path sums before they are done
    ops = schedule(4, 0, 2, pair=True)
    rec = ("record", "a", ("flushed", 0))
    ops.remove(rec)
    ops.insert(ops.index(("record", "s", ("shaded", 1))), rec)
    bad, _ = hazards(ops)
    assert any({a, b} == {"shadow finish 0.0", "shade 2.0"} and "acc[0]" in c for a, b, c in bad), bad[:5]


def test_model_catches_two_samples_in_flight_on_one_slot():
    # three samples in flight (waits on sample it - 3's flush) over only
    # two slots' buffers: samples 1 and 3 share slot 1 while sample 1's
    # shadow work may still run
    ops = schedule(6, 0, 3, slot_buffers=lambda it: it % 2)
    bad, _ = hazards(ops)
    assert bad and any("shq[1]" in c or "acc[1]" in c for _, _, c in bad)
    # and a slot whose histogram is never cleared is reported
    ops = [o for o in schedule(4, 0, 3) if not (o[0] == "op" and o[2] == "reset shist[2]")]
    _, unclear = hazards(ops)
    assert unclear == ["sort 2.0"]
