"""Host-side load-step control shared by the plain Newton paths: the
adaptive sub-increment stack of the reference's adaptive solution
(adapt_check.f, called from mnralg.f).

A failing increment is split into 4 sub-increments, recursively, down to
1/16 of the step; the rows of the stack are (start, end, increment,
extrapolation scale) with the current sub-increment last.  The first
sub-increment after a split restarts without extrapolation; the later
ones extrapolate the last converged sub-increment by the row's scale.
"""


class Adaptive:
    """adapt_check.f: rows (start, end, increment, extrapolation scale) of
    the sub-increments still to run, the current one last."""
    divisions = 4

    def __init__(self):                       # process type 3
        self.stack = [[0.0, 1.0, 1.0, 1.0]]
        self.min_fact = 1.0 / self.divisions ** 2
        self.scaling = 1.0
        self.disp_fact = 1.0
        self.load_fact = 1.0

    def cut(self):                            # process type 1
        """Split the current increment; False when it is already the
        smallest allowed."""
        if self.disp_fact <= self.min_fact:
            return False
        st = self.stack
        lv = len(st) - 1
        a, b = st[lv][0], st[lv][1]
        n = self.divisions
        inc = (b - a) / n
        rows = []
        start = a
        for _ in range(n):                    # pushed last-first
            rows.append([start, start + inc, inc, 1.0])
            start = start + inc
        rows.reverse()
        st.extend(rows)
        top = len(st) - 1
        st[top][3] = st[lv][3] / n
        st[lv][3] = st[lv][3] / st[top][3]
        for k in range(lv + 1, top):
            st[k][3] = 1.0
        self.scaling = st[top][3]
        self.disp_fact = st[top][2]
        self.load_fact = st[top][1]
        return True

    def advance(self):                        # process type 2
        """After a converged sub-increment: True if another one follows."""
        st = self.stack
        if len(st) <= 1:
            return False
        lv = len(st) - 1
        if st[lv][1] == st[lv - 1][1]:
            first_row = st[0][3]
            del st[lv - 1:]
            if not st:
                self.stack = [[0.0, 1.0, 1.0, first_row]]
                self.scaling = first_row
                return False
            lv = len(st) - 1
            st[lv][3] = self.divisions * st[lv][3]
            self.scaling = st[lv][3]
        else:
            del st[lv:]
        lv = len(st) - 1
        if lv <= 0:
            return False
        self.scaling = st[lv][3]
        self.disp_fact = st[lv][2]
        self.load_fact = st[lv][1]
        return True
