"""Host-side logic of the permanent's chunk-shape autotuner: the choice rule on
measured times, and the table rewrite (no GPU)."""
import re

from gbskernels_amd import autotune


def test_choice_keeps_the_entry_unless_a_candidate_is_clearly_better():
    # batch times (ms) and lone-matrix times (ms) per candidate k
    tb = {7: 10.0, 8: 9.9, 9: 9.85, 10: 9.8}
    t1 = {7: 0.05, 8: 0.08, 9: 0.12, 10: 0.30}
    # every batch time within 2% of the best and every lone time under the
    # 0.25 ms floor except k = 10: the entry (9) stays
    assert autotune.choose(tb, t1, cur=9, ref=9, latency_ms=0.25, margin=0.02) == 9
    # a 5% faster batch time wins when its lone matrix stays admissible
    tb2 = {**tb, 10: 9.3}
    assert autotune.choose(tb2, {**t1, 10: 0.2}, cur=9, ref=9, latency_ms=0.25, margin=0.02) == 10
    # ... but not when the lone matrix gets 2.5x slower than the default's
    assert autotune.choose(tb2, t1, cur=9, ref=9, latency_ms=0.1, margin=0.02) == 9


def test_choice_prefers_the_faster_lone_matrix_within_the_batch_band():
    tb = {6: 60.0, 9: 50.2, 10: 50.0, 11: 50.1, 12: 50.3}
    t1 = {6: 15.5, 9: 10.4, 10: 10.2, 11: 10.0, 12: 19.8}
    # current entry 12: inside the batch band but its lone matrix is ~2x the
    # best one in the band, so the tuner moves to 11
    assert autotune.choose(tb, t1, cur=12, ref=6, latency_ms=0.25, margin=0.02) == 11
    # the default is outside the band (20% slower batch): replaced as well
    assert autotune.choose(tb, t1, cur=6, ref=6, latency_ms=0.25, margin=0.02) == 11


def test_write_table_rewrites_only_the_given_entries(tmp_path):
    src = autotune.TUNE_HEADER.read_text()
    hdr = tmp_path / "perm_tune.hpp"
    hdr.write_text(src)
    out = tmp_path / "out.hpp"
    autotune.write_table({0: {24: 12}, 1: {20: 8}}, path=hdr, out=out)
    before = [[int(x) for x in r.split(",")] for r in re.findall(r"\{([0-9, ]+)\}", src)]
    after = [[int(x) for x in r.split(",")] for r in re.findall(r"\{([0-9, ]+)\}", out.read_text())]
    assert len(after) == 2 and all(len(r) == 29 for r in after)
    assert after[0][24] == 12 and after[1][20] == 8
    for fam, (b, a) in enumerate(zip(before, after)):
        changed = [n for n in range(29) if a[n] != b[n]]
        assert set(changed) <= {24 if fam == 0 else 20}
    assert hdr.read_text() == src  # the source table is untouched


def test_batches_cover_the_config_sizes():
    assert autotune.batch_for(16, 0) == 65536 and autotune.batch_for(24, 0) == 1024
    assert autotune.batch_for(28, 0) == 64 and autotune.batch_for(28, 1) == 16


def test_rewrite_constant_touches_only_the_named_built_in(tmp_path):
    """--tor / --haf --write rewrite one integer constant each, in place."""
    from gbskernels_amd import autotune
    for src, names in ((autotune.TOR_SOURCE, ("TOR_BFS_MAX_F64", "TOR_BFS_MAX_DD", "TOR_GROUPED_MAX_F64",
                                              "TOR_GROUPED_MAX_DD")),
                       (autotune.HAF_SOURCE, ("haf_small_max_n", "haf_small_target_wgs"))):
        out = tmp_path / src.name
        out.write_text(src.read_text())
        for k, name in enumerate(names):
            autotune._rewrite_constant(src, name, 7 + k, out)
        text = out.read_text()
        for k, name in enumerate(names):
            assert f"{name} = {7 + k};" in text
        changed = [a for a, b in zip(src.read_text().splitlines(), text.splitlines()) if a != b]
        assert len(changed) == len(names)
