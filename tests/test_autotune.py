"""Launch-shape parameters: the C ABI's names, ranges and per-box override,
and the shipped tuning table's shape (no GPU: the context entry points
need a device, so only the table and the host helpers run here)."""
import json

from emclarity_amd import autotune


def test_shipped_table_lists_candidates_it_chose():
    t = autotune.load_table()
    if not t:
        return                      # no table shipped: library defaults apply
    for table, by_n in t["params"].items():
        assert table in ("c64", "f16")
        for n, p in by_n.items():
            assert int(n) in (64, 96, 128)
            for name, v in p.items():
                assert v in autotune.CANDIDATES[name]
                times = t["timings_ms"][table][n][name]
                assert float(times[str(v)]) == min(float(x) for x in times.values())


def test_params_for_and_empty_override(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"params": {"c64": {"64": {"search_tile": 8}}}}))
    t = autotune.load_table(p)
    assert autotune.params_for(64, "c64", t) == {"search_tile": 8}
    assert autotune.params_for(96, "c64", t) == {} and autotune.params_for(64, "f16", t) == {}
