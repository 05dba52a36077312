# Timing-only edit of SOURCES/update.f90 for the CPU baseline: in entropy_residual the reference
# assigns the whole array `rescale` once per node inside a DO loop over the nodes
# (SOURCES/update.f90:865-867), O(np^2) work per stage that gives the same array every time.  This
# filter drops that DO line and its END DO and keeps the assignment, so the edited build computes
# the same bits in O(np).  Every other line passes through unchanged; oracle/Makefile checks that
# exactly two lines were removed.
{
    if (state == 1) {
        if ($0 ~ /^ *rescale = ABS\(maxn-minn\)/) { print; state = 2; next }
        print held
        state = 0
    } else if (state == 2) {
        state = 0
        if ($0 ~ /^ *END DO *$/) next
    }
    if (state == 0 && $0 ~ /^ *DO i = 1, mesh%np *$/) { held = $0; state = 1; next }
    print
}
END { if (state == 1) print held }
