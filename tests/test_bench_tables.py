"""The bench's rocprofv3 table logic on small synthetic databases: counter
scaling by the copy calibration, the shade stage counted per bounce (all class
kernels of a bounce summed), and which kernel instances make a stage."""
import os
import sqlite3
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, dispatch_id integer, value real, "
                "counter_name text, duration real)")
    con.executemany("insert into counters_collection values (?, ?, ?, 'X', 1000.0)", rows)
    con.commit()
    con.close()
    return str(path)


def test_stage_instances():
    assert bench._stage_of("rt::k_extend<false, false>") == "k_extend"
    assert bench._stage_of("rt::k_extend<true, false>") is None       # counting render
    assert bench._stage_of("rt::k_extend<false, true>") is None       # pooled
    assert bench._stage_of("rt::k_extend<false, false, 16>") == "k_extend"  # with straggler suspension
    assert bench._stage_of("rt::k_extend_cont<false, false>") == "k_extend"
    assert bench._stage_of("rt::k_extend_cont<true, false>") is None
    assert bench._stage_of("rt::k_shadow<false, false, 0, 16>") == "k_shadow"
    assert bench._stage_of("rt::k_shadow_cont<false, false>") == "k_shadow"
    assert bench._stage_of("rt::k_shadow<false, false, 2, 16>") is None
    assert bench._stage_of("rt::k_shadow<false, false, 0>") == "k_shadow"
    assert bench._stage_of("rt::k_shadow<false, false, 2>") is None   # another shadow test
    assert bench._stage_of("rt::k_shade<false>") == "k_shade"
    assert bench._stage_of("rt::k_shade<1, false>") == "k_shade"
    assert bench._kernel_name("void rt::(anonymous namespace)::k_bvh_node<1>(int, float*)") == "rt::k_bvh_node<1>"


def test_traffic_per_bounce_and_scaling(tmp_path):
    # Two renders of 3 bounces: 6 plain extend dispatches, plus one counting render of
    # 3 bounces whose shade launches are the same kernel.  Shade has one launch per bounce.
    ext = [("void rt::k_extend<false, false>(args)", d, 100.0) for d in range(6)]
    ext += [("void rt::k_extend<true, false>(args)", 10 + d, 500.0) for d in range(3)]
    shade = [("void rt::k_shade<false>(args)", 20 + d, 40.0) for d in range(9)]
    shadow = [("void rt::k_shadow<false, false, 0>(args)", 40 + d, 60.0) for d in range(6)]
    fetch = _db(tmp_path / "f.db", ext + shade + shadow)
    write = _db(tmp_path / "w.db", [(k, d, v / 10) for k, d, v in ext + shade + shadow])
    # A render whose extend launches suspend stragglers: each bounce is a k_extend
    # launch plus its continuation launch, counted together per bounce.
    ext2 = [("void rt::k_extend<false, false, 16>(args)", d, 80.0) for d in range(6)]
    ext2 += [("void rt::k_extend_cont<false, false>(args)", 50 + d, 20.0) for d in range(6)]
    ext2 += [("void rt::k_extend<true, false, 16>(args)", 10 + d, 500.0) for d in range(3)]
    shadow2 = [("void rt::k_shadow<false, false, 0, 16>(args)", 40 + d, 45.0) for d in range(6)]
    shadow2 += [("void rt::k_shadow_cont<false, false>(args)", 60 + d, 15.0) for d in range(6)]
    fetch2 = _db(tmp_path / "f2.db", ext2 + shade + shadow2)
    write2 = _db(tmp_path / "w2.db", [(k, d, v / 10) for k, d, v in ext2 + shade + shadow2])
    gib = 1 << 30
    # The copy kernel reads half its bytes on FETCH_SIZE and all of them on WRITE_SIZE.
    cf = _db(tmp_path / "cf.db", [("__amd_rocclr_copyBuffer", 1, gib / 1024 / 2)])
    cw = _db(tmp_path / "cw.db", [("__amd_rocclr_copyBuffer", 1, gib / 1024)])
    out = bench.traffic_from_pmc(fetch, write, cf, cw)
    rows = {ln.split()[0]: ln.split() for ln in out.splitlines() if ln.startswith("k_")}
    assert out.splitlines()[0] == "fetch_scale 2.0000"
    # extend: the plain dispatches only, 100 KiB x 2 fetched and 10 KiB written per launch
    assert rows["k_extend"][1:] == ["6", str(200 * 1024), str(10 * 1024), str(210 * 1024)]
    # shade: 9 dispatches over 9 bounces (6 plain + 3 counting)
    assert rows["k_shade"][1:] == ["9", str(80 * 1024), str(4 * 1024), str(84 * 1024)]
    assert rows["k_shadow"][1:] == ["6", str(120 * 1024), str(6 * 1024), str(126 * 1024)]
    out2 = bench.traffic_from_pmc(fetch2, write2, cf, cw)
    rows2 = {ln.split()[0]: ln.split() for ln in out2.splitlines() if ln.startswith("k_")}
    assert rows2["k_extend"][1:] == rows["k_extend"][1:]
    assert rows2["k_shade"][1:] == rows["k_shade"][1:]
    assert rows2["k_shadow"][1:] == rows["k_shadow"][1:]
    (tmp_path / "traffic.txt").write_text("# header\n" + out + "\n")
    assert bench.load_traffic(str(tmp_path / "traffic.txt")) == {
        "k_extend": 210 * 1024, "k_shade": 84 * 1024, "k_shadow": 126 * 1024}
    assert bench.load_traffic(str(tmp_path / "absent.txt")) == {}


def test_burley_variant_rewrites_only_the_method(tmp_path):
    import json

    from cudaraytracer_amd import scenes

    path = scenes.write_scene(scenes.terrain_env(grid=11, env_w=8, env_h=4, width=8, height=8, spp=1),
                              str(tmp_path), "t")
    out = bench.burley_variant(path)
    a, b = json.load(open(path)), json.load(open(out))
    assert os.path.dirname(out) == os.path.dirname(path)  # same OBJ and map paths
    assert [m.get("subsurface_method") for m in a["materials"]] == [None, "random_walk"]
    assert [m.get("subsurface_method") for m in b["materials"]] == [None, "burley"]
    for m in b["materials"]:
        m.pop("subsurface_method", None)
    for m in a["materials"]:
        m.pop("subsurface_method", None)
    assert a == b
