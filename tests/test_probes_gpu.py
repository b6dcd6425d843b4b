"""Measurement probes (include/hlanc_probes.h): each returns a positive time on
its own buffers; the grid-barrier probe refuses a grid that cannot be
co-resident instead of launching it, and its barriers carry data (a block reads
its neighbour's writes of the same round)."""
import pytest

import hlanc_amd as H

pytestmark = pytest.mark.gpu


def test_stream_and_gather_rates_are_plausible():
    t = H.bench_stream(1 << 30, reps=5)
    assert 1.0e9 / t > 1.0e12  # > 1 TB/s of reads
    assert H.bench_gather(1 << 16, 1 << 24, random=True, reps=3) > 0.0


@pytest.mark.parametrize("blocks", [8, 64, 256])
def test_grid_barrier_probe(blocks):
    t = H.bench_grid_barrier(blocks, barriers=200)
    assert 0.0 < t < 1e-3


def test_grid_barrier_probe_refuses_a_grid_that_cannot_be_resident():
    with pytest.raises(H.HLancError):
        H.bench_grid_barrier(1 << 20, barriers=1)


def test_graph_node_probe():
    t = H.bench_graph_nodes(nodes=20, blocks=256, reps=20)
    assert 0.0 < t < 1e-4
