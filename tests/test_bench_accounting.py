"""bench.py's algorithmic-byte formula of the tiled raster, run on the CPU.

tile_algorithmic_bytes reads the bins the tiled raster left in a view's
scratch (work items, (draw, part) entries, bin headers) and prices each entry
at 12 bytes per cached vertex, 4 per triangle of its part and 12 for the entry
and header words, plus 8 per covered pixel.  Here the scratch holds three
hand-built entries in host memory: a whole meshlet, the second half of a
meshlet split in two, and another whole meshlet."""
import sys
from pathlib import Path

import numpy as np

from lighthugger_amd import abi
from lighthugger_amd.frame import HostMemory, SceneData, ViewConfig, ViewData
from lighthugger_amd.scene import HostScene

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def test_tile_algorithmic_bytes_prices_each_entry_of_the_bins():
    import bench

    scene = HostScene.preset(abi.PRESET_SMALL, tri_scale_pct=25, texture_size=16)
    mem = HostMemory()
    view = ViewData(ViewConfig(128, 64, 64, bin_capacity=64), SceneData(scene, mem), mem)
    cfg = view.cfg
    L = abi.scratch_layout(cfg.width, cfg.height, view.dcap, cfg.big_capacity, cfg.clip_capacity, view.bcap)
    scratch = view.offsets["scratch"]

    draws = np.zeros(3, abi.MESHLET_REFERENCE)
    draws["instance_index"] = [0, 0, 1]
    draws["meshlet_index"] = [0, 1, 0]
    mem.write(view.handle, view.offsets["draws"], draws)
    counters = np.zeros(abi.NUM_COUNTERS, np.uint32)
    counters[abi.CNT_DRAWS] = 3
    counters[abi.CNT_TILE_ITEMS] = 2
    mem.write(view.handle, view.offsets["counters"], counters)
    # items (tile, begin, end | only): tile 0 holds entries 0-1, tile 1 entry 2 alone
    items = np.array([[0, 0, 2], [1, 2, 3 | 0x80000000]], np.uint32)
    mem.write(view.handle, scratch + L["items"], items)
    # entries: draw | part << 25
    entries = np.array([0, 1 | (1 << 25), 2], np.uint32)
    mem.write(view.handle, scratch + L["entries"], entries)
    # headers (rect, ntri | log2 parts << 8, mid, 0): 10 triangles; 9 in 2 parts; 5
    hdr = np.zeros((3, 4), np.uint32)
    hdr[:, 1] = [10, 9 | (1 << 8), 5]
    mem.write(view.handle, scratch + L["hdr"], hdr)

    def vertices(instance, meshlet):
        mi = int(scene.instances["mesh_info_address"][instance])
        return int(scene.meshlets(mi)["index_count"][meshlet])

    nv = [vertices(0, 0), vertices(0, 1), vertices(1, 0)]
    covered = 1000
    # part 1 of 2 of 9 triangles: triangles [9 // 2, 9) = 5 of them
    expected = (12 * nv[0] + 4 * 10 + 12) + (12 * nv[1] + 4 * 5 + 12) + (12 * nv[2] + 4 * 5 + 12) + 8 * covered
    got = bench.tile_algorithmic_bytes(scene, view, bench.meshlet_sizes(scene), covered)
    assert got == expected
    # no work items: only the covered pixels' keys
    counters[abi.CNT_TILE_ITEMS] = 0
    mem.write(view.handle, view.offsets["counters"], counters)
    assert bench.tile_algorithmic_bytes(scene, view, bench.meshlet_sizes(scene), covered) == 8 * covered
