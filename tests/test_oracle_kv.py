"""The oracle's sorted key-value output pinned by what the reference's rules
fix without its code (geometry.cu:665-1107,1752-2138; step_merge.cpp:140-310):

- key layout and meeting boxes: every meeting the (separately pinned)
  meeting rule emits between two flushes appears as exactly two directed
  keys [sid|oid|target|duration|end-end_time_min] whose box is the float
  min/max of both objects' positions over the run, recomputed from the trace;
- oversize rule: keys of objects with a meeting longer than max_meet_time or
  an object box (float_to_uint/uint_to_float round trip) above the area limit
  go to the oversize buffer; counts follow the reference's accounting;
- total order of the unique 128-bit keys;
- STR: slab and part structure (sizes x_cap and x_cap/split+1, monotone in
  the mid's high then low word);
- files: capacity = keys of their objects, key ranges, limits, widths from
  int.bit_length (an independent floor(log2)+1), bitmap coverage and the
  packed stream decoded back to the keys' fields.
"""
import numpy as np
import pytest

from tests.kvcases import KvCase, fields, f2u, run_oracle, u2f

BITMAP = 8192


def expected_keys(run, case, k):
    """Directed keys of the meetings emitted in flush period k, with boxes
    recomputed from the trace."""
    lo_t = -1 if k == 0 else case.flush_at[k - 1]
    hi_t = case.flush_at[k]
    etmin = case.min_meet if k == 0 else case.flush_at[k - 1]   # start_time 0
    tr32 = run.trace.astype(np.float32)
    out = {}
    for t in range(lo_t + 1, hi_t + 1):
        for a, b, s, e in run.meetings[t].astype(np.int64):
            pts = np.concatenate([tr32[s:e + 1, a], tr32[s:e + 1, b]])
            box = np.array([pts[:, 0].min(), pts[:, 1].min(), pts[:, 0].max(), pts[:, 1].max()], np.float32)
            for p, q in ((a, b), (b, a)):
                out[(p, q, e - s, e - etmin)] = box
    return out


def all_keys(fl):
    keys = np.concatenate([fl["keys"], fl["okeys"]])
    boxes = np.concatenate([fl["boxes"], fl["oboxes"]])
    return keys, boxes


def object_aggregates(exp, n):
    cnt = np.zeros(n, np.int64)
    ob = np.zeros((n, 4), np.uint32)
    ob[:, 0] = ob[:, 1] = 100000000
    for (p, q, d, e), box in exp.items():
        cnt[p] += 1
        u = f2u(box)
        ob[p, 0] = min(ob[p, 0], u[0])
        ob[p, 1] = min(ob[p, 1], u[1])
        ob[p, 2] = max(ob[p, 2], u[2])
        ob[p, 3] = max(ob[p, 3], u[3])
    return cnt, ob


@pytest.fixture(scope="module")
def plain():
    case = KvCase(max_meet=4096, area_limit=1e9)
    return case, run_oracle(case)


@pytest.fixture(scope="module")
def cut():
    # short max_meet and a small area limit exercise both oversize routes
    case = KvCase(max_meet=9, area_limit=600.0, seed=7)
    return case, run_oracle(case)


@pytest.fixture(scope="module")
def wide():
    # a 1200 m box: float_to_uint of x > 820 m passes the object boxes'
    # initial low (1e8), so the lows saturate and the +180 offset shows
    case = KvCase(n=6000, side=1200.0, area_limit=3000.0, seed=11)
    return case, run_oracle(case)


CASES = ["plain", "cut", "wide"]


@pytest.mark.parametrize("which", CASES)
def test_keys_and_boxes_from_meetings(which, request):
    case, run = request.getfixturevalue(which)
    for k, fl in enumerate(run.flushes):
        exp = expected_keys(run, case, k)
        keys, boxes = all_keys(fl)
        sid, oid, tgt, dur, end = fields(keys)
        got = {(int(o), int(t), int(d), int(e)): boxes[i] for i, (o, t, d, e) in enumerate(zip(oid, tgt, dur, end))}
        assert len(got) == len(keys) == len(exp) > 20
        assert set(got) == set(exp)
        for key, box in exp.items():
            assert np.array_equal(got[key], box), key
        assert fl["summary"]["logged_meetings"] * 2 == len(exp)


@pytest.mark.parametrize("which", CASES)
def test_oversize_rule_and_counts(which, request):
    case, run = request.getfixturevalue(which)
    for k, fl in enumerate(run.flushes):
        exp = expected_keys(run, case, k)
        cnt, ob = object_aggregates(exp, case.n)
        l0, l1, h0, h1 = u2f(ob[:, 0]), u2f(ob[:, 1]), u2f(ob[:, 2]), u2f(ob[:, 3])
        area = ((h1 - l1) * (h0 - l0)).astype(np.float32)
        by_area = (cnt >= 1) & (area.astype(np.float64) > case.area_limit)
        long_meet = {(p, q) for (p, q, d, e) in exp if d > case.max_meet}
        oversize = by_area.copy()
        for p, q in long_meet:
            oversize[p] = oversize[q] = True
        _, koid, *_ = fields(fl["keys"])
        _, ooid, *_ = fields(fl["okeys"])
        assert not oversize[koid].any() and oversize[ooid].all()
        s = fl["summary"]
        assert s["sid_count"] == int((cnt >= 1).sum())
        assert s["oversize_oid_count"] == len(long_meet) // 2 + int(by_area.sum())
        assert s["zero_one_count"] == case.n - s["sid_count"] + s["oversize_oid_count"]
        if which == "cut":
            assert by_area.any() and long_meet
        if which == "wide":
            assert by_area.any() and (ob[cnt >= 1, 0] == 100000000).any()


@pytest.mark.parametrize("which", CASES)
def test_total_order_and_sid_field(which, request):
    case, run = request.getfixturevalue(which)
    for fl in run.flushes:
        for part in ("keys", "okeys"):
            k = fl[part]
            v = [(int(h) << 64) | int(l) for l, h in k]
            assert all(a < b for a, b in zip(v, v[1:]))
        sid, oid, *_ = fields(fl["keys"])
        assert np.array_equal(sid, fl["sids"][oid].astype(np.int64))
        osid, *_ = fields(fl["okeys"])
        assert (osid == 0).all()


def test_str_structure(plain):
    case, run = plain
    split = case.split
    for k, fl in enumerate(run.flushes):
        exp = expected_keys(run, case, k)
        cnt, ob = object_aggregates(exp, case.n)
        sids = fl["sids"].astype(np.int64)
        present = cnt >= 1
        assert (sids[~present] == 0).all() and (sids[present] >= 2).all()
        # mid = (low0/2 + high0/2) << 32 | (high1/2 + high1/2) (geometry.cu:763-765)
        mx = (ob[:, 0] // 2 + ob[:, 2] // 2).astype(np.int64)
        my = (ob[:, 3] // 2 + ob[:, 3] // 2).astype(np.int64)
        m = int(present.sum())
        xcap = m // split + 1
        pcap = xcap // split + 1
        slab = (sids - 2) // split
        part = (sids - 2) % split
        for i in range(split):
            members = np.flatnonzero(present & (slab == i))
            assert len(members) == min(xcap, max(0, m - i * xcap))
            if i + 1 < split:
                nxt = np.flatnonzero(present & (slab == i + 1))
                if len(members) and len(nxt):
                    assert (mx[members] << 32 | my[members]).max() <= (mx[nxt] << 32 | my[nxt]).min()
            for p in range(split):
                pm = members[part[members] == p]
                assert len(pm) == min(pcap, max(0, len(members) - p * pcap))
                nx = members[part[members] == p + 1]
                if len(pm) and len(nx):
                    assert my[pm].max() <= my[nx].min()


def min_bits(v: int) -> int:
    return max(1, int(v).bit_length())


@pytest.mark.parametrize("which", CASES)
def test_files(which, request):
    case, run = request.getfixturevalue(which)
    box = np.array([0, 0, case.side, case.side], np.float64)
    for k, fl in enumerate(run.flushes):
        keys, boxes, ctfs = fl["keys"], fl["boxes"], fl["ctfs"]
        sid, oid, tgt, dur, end = fields(keys)
        etmin = fl["summary"]["end_time_min"]
        assert len(ctfs) == case.split ** 2
        off = 0
        poff = 0
        for t, f in enumerate(ctfs):
            assert f["key_offset"] == off and f["packed_offset"] == poff
            n = int(f["kv_count"])
            rng = slice(off, off + n)
            if which == "plain":
                assert (sid[rng] == t + 2).all()
                assert n == int((fl["sids"][oid] == t + 2).sum())
            off += n
            poff += n * (int(f["key_bit"]) // 8)
            if n == 0:
                assert f["key_bit"] == 0
                continue
            b = boxes[rng]
            mbr = np.array([b[:, 0].min(), b[:, 1].min(), b[:, 2].max(), b[:, 3].max()], np.float32)
            assert np.array_equal(f["mbr"], mbr)
            starts = end[rng] + etmin - dur[rng]
            assert (f["start_time_min"], f["start_time_max"]) == (starts.min(), starts.max())
            assert (f["end_time_min"], f["end_time_max"]) == (etmin, case.flush_at[k])
            # widths (step_merge.cpp:145-186)
            wx = float(np.float32(mbr[2] - mbr[0]))
            wy = float(np.float32(mbr[3] - mbr[1]))
            lxg = int(wx / ((box[2] - box[0]) / 3000))
            lyg = int(wy / ((box[3] - box[1]) / 4200))
            lxb, lyb = min_bits(max(lxg, 1) - 1), min_bits(max(lyg, 1) - 1)
            eb = max(lxb, lyb)
            assert (f["id_bit"], f["duration_bit"], f["end_bit"]) == (
                min_bits(case.n), min_bits(case.max_meet - 1), min_bits(case.max_meet - 1))
            assert (f["low_x_bit"], f["low_y_bit"], f["edge_bit"], f["mbr_bit"]) == (lxb, lyb, eb, lxb + lyb + 2 * eb)
            total = 2 * int(f["id_bit"]) + 2 * int(f["duration_bit"]) + int(f["mbr_bit"])
            assert int(f["key_bit"]) == -(-total // 8) * 8
            ow, oh = int(wx / (box[2] - box[0]) * 300), int(wy / (box[3] - box[1]) * 420)
            xg, yg = max(ow + 1, 32), max(oh + 1, 32)
            if xg >= 256 or yg >= 256:
                xg, yg = min(ow + 1, 255), min(oh + 1, 255)
            assert (f["x_grid"], f["y_grid"], f["bitmap_bytes"]) == (xg, yg, (xg * yg + 7) // 8)
        assert poff == fl["summary"]["packed_bytes"]


@pytest.mark.parametrize("which", CASES)
def test_bitmaps_cover_the_key_boxes(which, request):
    case, run = request.getfixturevalue(which)
    for fl in run.flushes:
        keys, boxes, ctfs, bm = fl["keys"], fl["boxes"], fl["ctfs"], fl["bitmaps"]
        for t, f in enumerate(ctfs):
            got = np.unpackbits(bm[t * BITMAP:(t + 1) * BITMAP], bitorder="little")
            want = np.zeros(BITMAP * 8, np.uint8)
            n = int(f["kv_count"])
            if n and f["key_bit"]:
                xg, yg = int(f["x_grid"]), int(f["y_grid"])
                m = f["mbr"]
                w, h = np.float32(m[2] - m[0]), np.float32(m[3] - m[1])
                for b in boxes[f["key_offset"]:f["key_offset"] + n]:
                    q = lambda v, o, d, g: int(np.float32(np.float32(v - o) / d if d else 0) * np.float32(g))
                    x0, x1 = q(b[0], m[0], w, xg), q(b[2], m[0], w, xg)
                    y0, y1 = q(b[1], m[1], h, yg), q(b[3], m[1], h, yg)
                    for x in range(x0, min(x1, xg - 1) + 1):
                        for y in range(y0, min(y1, yg - 1) + 1):
                            want[x + y * xg] = 1
            assert np.array_equal(got, want), t
        # the oversize map bitmap: 256x256 cells of the root box
        want = np.zeros(BITMAP * 8, np.uint8)
        for b in fl["oboxes"]:
            c = [int((float(v) - 0.0) / case.side * 256) for v in b]
            for x in range(c[0], min(c[2], 255) + 1):
                for y in range(c[1], min(c[3], 255) + 1):
                    want[x + y * 256] = 1
        assert np.array_equal(np.unpackbits(fl["obitmap"], bitorder="little"), want)


@pytest.mark.parametrize("which", CASES)
def test_packed_stream_decodes_to_the_keys(which, request):
    case, run = request.getfixturevalue(which)
    for fl in run.flushes:
        keys, boxes, ctfs, packed = fl["keys"], fl["boxes"], fl["ctfs"], fl["packed"]
        sid, oid, tgt, dur, end = fields(keys)
        for f in ctfs:
            n, kb = int(f["kv_count"]), int(f["key_bit"]) // 8
            if not n or not kb:
                continue
            ib, db, eb_, mb = int(f["id_bit"]), int(f["duration_bit"]), int(f["end_bit"]), int(f["mbr_bit"])
            s_pid = ib + db + eb_ + mb
            present = max(0, min(n, len(keys) - int(f["key_offset"])))   # ranges clip at the kept keys
            for j in range(present):
                i = int(f["key_offset"]) + j
                o = int(f["packed_offset"]) + j * kb
                v = int.from_bytes(packed[o:o + kb].tobytes(), "little")
                # the key is a sum of shifted fields (the pid keeps the sid bits
                # above the oid, extract_fields), truncated to key_bit/8 bytes
                full_pid = (int(sid[i]) << 26) | int(oid[i])
                fields_sum = (full_pid << s_pid) + (int(tgt[i]) << (db + eb_ + mb)) + \
                    (int(dur[i]) << (eb_ + mb)) + (int(end[i]) << mb)
                code = (v - fields_sum) % (1 << (8 * kb))
                assert code < (1 << mb)
                lx, ly, e = int(f["low_x_bit"]), int(f["low_y_bit"]), int(f["edge_bit"])
                assert code >> (ly + 2 * e) < (1 << lx)
                if end[i] < (1 << eb_) and dur[i] < (1 << db):
                    # fields that fit their widths decode back
                    assert v >> mb & ((1 << eb_) - 1) == end[i]
                    assert v >> (mb + eb_) & ((1 << db) - 1) == dur[i]
                    assert v >> (mb + eb_ + db) & ((1 << ib) - 1) == tgt[i]
                    assert v >> s_pid == full_pid & ((1 << (8 * kb - s_pid)) - 1)


def test_two_flushes_restart_the_end_offsets(plain):
    case, run = plain
    a, b = run.flushes
    assert a["summary"]["end_time_min"] == case.min_meet
    assert b["summary"]["end_time_min"] == case.flush_at[0]
    assert b["summary"]["end_time_max"] == case.flush_at[1]
