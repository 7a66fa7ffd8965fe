"""The VSM ring layout (openfawe_amd.vsm.ring_layout, host code of the product) against the reference's definition
(vsm_wake.py:177-185: pack_rings with a unit strength on one panel, zero on the others), and the oracle of the
material edge cores (cores.py:129-149) on a hand-derived lattice.  CPU only."""
import numpy as np
import pytest

from oracle import vsm_np as VO
from openfawe_amd import vsm


PANEL_SETS = {
    "sequential-8": np.column_stack([np.arange(8), np.arange(1, 9)]),
    "scrambled-shared-root": np.array([[0, 1], [2, 1], [2, 3], [4, 5], [6, 5], [6, 7]]),
    "reversed": np.column_stack([np.arange(1, 6), np.arange(5)]),
}


@pytest.mark.parametrize("name", list(PANEL_SETS))
@pytest.mark.parametrize("depth", [2, 3, 5])
@pytest.mark.parametrize("at_load_points", [False, True])
def test_ring_layout_is_the_per_panel_pack_rings(name, depth, at_load_points):
    panels = PANEL_SETS[name]
    width = int(panels.max()) + 1
    segments, pointer, column_segment, column_sign = vsm.ring_layout(depth, width, panels, at_load_points)
    columns = VO.ring_columns(depth, width, panels, at_load_points)
    assert len(pointer) == len(panels) + 1 and pointer[0] == 0
    for j, (edges, gamma) in enumerate(columns):
        k = slice(pointer[j], pointer[j + 1])
        np.testing.assert_array_equal(segments[column_segment[k]], edges)
        np.testing.assert_array_equal(column_sign[k], gamma)
    # segments are numbered by first use along the columns, each used
    first_use = [np.flatnonzero(column_segment == s)[0] for s in range(len(segments))]
    assert first_use == sorted(first_use)
    if at_load_points:
        assert not np.any((segments[:, 0] < width) & (segments[:, 1] < width))


def test_edge_radii_known_answer():
    # width 3, panels (0,1), (1,2); rows 0 (bound), 1 (trailing edge), 2 and 3 (free)
    panels = np.array([[0, 1], [1, 2]])
    span = np.array([[10.0, 11.0], [12.0, 13.0]])               # rows 2 and 3
    stream = np.array([[20.0, 21.0, 22.0], [23.0, 24.0, 25.0], [26.0, 27.0, 28.0]])
    edges = np.array([[0, 1], [3, 4], [7, 8], [10, 11], [2, 5], [4, 7], [6, 9], [1, 0]])
    got = VO.edge_radii(edges, 3, panels, 0.5, span, stream)
    # [7,8]: row 2, columns 1-2 -> panel 1 -> span[0, 1]; [10,11]: row 3 -> span[1, 1]; [4,7]: rows 1-2 at
    # column 1 -> stream[0, 1]; [6,9]: rows 2-3 at column 0 -> stream[1, 0]; bound and trailing-edge edges keep 0.5
    assert got.tolist() == [0.5, 0.5, 11.0, 13.0, 0.5, 21.0, 23.0, 0.5]
    assert VO.edge_radii(edges, 3, panels, 0.5).tolist() == [0.5] * 8
    with pytest.raises(ValueError):
        VO.edge_radii(np.array([[0, 4]]), 3, panels, 0.5, span, stream)
