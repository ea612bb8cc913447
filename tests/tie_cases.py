"""Co-optimal alignments whose placement is fixed only by the aligners'
tie-break rules, each worked by hand on the full DP matrix from the
reference's lines (scores 5 / -4 / -8, end gaps free, band 16 so the band
never binds at these lengths):

  vectorized (nwalign.rs:700-737 dploop, 993-1026 end-gap recalc, 1101-1139
  traceback and un-swap): s1 is the shorter input; without swap a tie goes
  up (3) >= left (2) >= diag (1), with swap left >= up >= diag; diag wins only
  when strictly greater.  On the last row (all of s1 used) a free left move
  replaces the cell when larger, or on a tie when the cell's pointer is diag
  (no swap) / not left (swap); on the last column a free up move replaces it
  when larger, or on a tie when the pointer is not up (no swap) / is diag
  (swap).  The last cell takes the left recalc first, then the right one.
  scalar ends-free and homopolymer (nwalign.rs:336-417, 435-528, traceback
  234-273): no swap; up if up >= diag and up >= left, else left if
  left >= diag, else diag; left moves are free on the last row and up moves
  on the last column; the homopolymer variant charges homo_gap_p for a gap
  opposite a base inside a run of >= 3 (nwalign.rs:211-227).

Each entry: (aligner, s1, s2, expected al0, expected al1, score, derivation).
"""

TIE_CASES = [
    # d(3,4): left = d(3,3) - 8 = 7 ties diag = d(2,3) + 5 = 7, diag needs
    # strictly more, so the gap in s1 takes the run's last A; traceback
    # (4,5) diag, (3,4) left, (3,3)..(1,1) diag.
    ("vectorized", "GAAC", "GAAAC", "GAA-C", "GAAAC", 12,
     "insertion placed at the run's last position (left beats diag on a tie)"),
    # swapped inputs: the DP runs on (GAAC, GAAAC) with left >= up; the path
    # has no up/left tie, so the same columns come back un-swapped.
    ("vectorized", "GAAAC", "GAAC", "GAAAC", "GAA-C", 12,
     "the same placement under swap"),
    # last cell (2,3): regular value diag 2; left recalc d(2,2) = 5 > 2 sets
    # left; right recalc d(1,3) = 5 ties it and, without swap, the pointer is
    # not up, so up wins: traceback up, diag (1,3), then row 0 lefts.
    ("vectorized", "AA", "ATA", "--AA", "ATA-", 5,
     "without swap the last-column free gap takes the tie"),
    # swapped: same matrix with left >= up; the right recalc's tie keeps the
    # left pointer (it is not diag): traceback left, left, diag (2,1), up.
    ("vectorized", "ATA", "AA", "-ATA", "AA--", 5,
     "with swap the last-row free gap keeps the tie"),
    # (1,2): diag 5 ties the last-row free gap d(1,1) + 0 = 5; the pointer was
    # diag, so the free gap takes it: the single A pairs with s2's first A.
    ("vectorized", "A", "AA", "A-", "AA", 5, "a free trailing gap beats diag on a tie"),
    ("vectorized", "AA", "A", "AA", "A-", 5, "the same under swap"),
    # scalar: d(3,4) left 7 >= diag 7 gives left; corner (4,5) diag 12.
    ("endsfree", "GAAC", "GAAAC", "GAA-C", "GAAAC", 12, "left beats diag on a tie"),
    # corner (2,3): free up d(1,3) = 5 and free left d(2,2) = 5 tie; up wins.
    ("endsfree", "AA", "ATA", "--AA", "ATA-", 5, "up beats left at the corner"),
    # s1 = ATA: corner (3,2) up 5 = left 5 > diag 2 -> up; (2,2) free up 5.
    ("endsfree", "ATA", "AA", "-ATA", "AA--", 5, "no swap in the scalar aligner"),
    # homopolymer: s2's AAA run makes a gap in s1 there cost -1; d(3,4):
    # left = d(3,3) - 1 = 14 ties diag = d(2,3) + 5 = 14, left wins; the
    # corner is diag 19 = 4 matches - 1.
    ("homo", "CAAG", "CAAAG", "CAA-G", "CAAAG", 19, "homopolymer gap at the run's end"),
]
