"""The packed SW walk reads its flag words from an LDS window of kWinC columns
x kWinQ words that one half-wave loads around the anchor cell (iw, jw)
(genoaligner_amd/csrc/sw16.hip, walk16).  These tests check, over every anchor
alignment and lane, that each word a walk step may read lies inside the loaded
window: the current cell in both trace layouts, and in the H-mod-4 layout also
the cell above and the cell to the left (a gap step reads them).  The window
constants and the word index are those of the kernel."""
import itertools

K_WIN_R, K_WIN_C, K_WIN_Q = 32, 32, 9


def _word_index(iw, jw, ii, jj):
    l = (jj - 1) & 63                       # lane of column jj within its strip
    t = ii - 1 + l                          # step at which that lane computed row ii
    return jw - jj, (t >> 2) - ((iw - K_WIN_R + l) >> 2)


def _outside(hmod, rows, cols):
    bad = 0
    for iw, jw in itertools.product(range(30, 42), range(2, 134)):   # all row alignments, all lanes
        for di, dj in itertools.product(range(rows), range(cols)):
            i, j = iw - di, jw - dj
            if i < 1 or j < 1:
                continue
            cells = [(i, j)] + ([(i, j - 1), (i - 1, j)] if hmod else [])
            for ii, jj in cells:
                if ii < 1 or jj < 1:
                    continue
                col, q = _word_index(iw, jw, ii, jj)
                bad += not (0 <= col < K_WIN_C and 0 <= q < K_WIN_Q)
    return bad


def test_flag_layout_steps_stay_in_the_window():
    assert _outside(False, K_WIN_R, K_WIN_C) == 0


def test_hmod_layout_needs_and_has_a_one_row_one_column_margin():
    assert _outside(True, K_WIN_R - 1, K_WIN_C - 1) == 0
    # without the row margin a gap step reads below the loaded words
    assert _outside(True, K_WIN_R, K_WIN_C - 1) > 0
