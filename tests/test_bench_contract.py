"""bench.py's host-side contract, without a GPU: the multi-GPU refusal, the
variant table of the Barnes-Hut block and the roofline figures it reads from the
counter summaries in profiles/."""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_variant_labels_unique_and_counter_keys_name_variants():
    labels = [v[0] for v in bench.bh_variants(True)]
    assert len(labels) == len(set(labels))
    for key, label in bench.BH_COUNTER_VARIANT.items():
        assert label in labels, (key, label)


def test_roofline_figures_follow_the_preferred_variant():
    n = 1 << 22
    morton = "morton tree: quadrupole theta=0.6, local far field, near 2"
    heap = "quadrupole theta=0.62, local far field, near 3"
    r = bench.bh_roofline_from_profile(n, morton)
    assert r["variant"] == morton and r["source"].endswith("bh_morton_two_level_walk_plummer_4194304.md")
    assert r["valu"] > 0 and r["traffic"] > 0
    r = bench.bh_roofline_from_profile(n, heap)
    assert r["variant"] == heap and r["source"].endswith("bh_two_level_walk_counters_plummer_4194304.md")
    # a variant without counters falls back to one that has them
    assert bench.bh_roofline_from_profile(n, "plain theta=0.5") is not None
    assert bench.bh_roofline_from_profile(1 << 10, morton) is None


def test_direct_traffic_figure_present():
    assert bench.traffic_from_profile(1 << 18) > 0


def test_roofline_falls_back_to_the_same_tree():
    n = 1 << 22
    r = bench.bh_roofline_from_profile(n, "morton tree: quadrupole theta=0.7, local far field, near 3")
    assert r["variant"].startswith("morton tree:")
    r = bench.bh_roofline_from_profile(n, "quadrupole theta=0.8, local far field, near 5")
    assert not r["variant"].startswith("morton tree:")
