"""Three literal restatements of the reference's beta <-> DFT maps (even sizes).

TEST INFRASTRUCTURE ONLY.  The reference states the same packing three times,
independently: a per-output-index gather form, a loop form and an array
concatenation form.  They are restated here one by one (1-based indices kept as
in the reference so each line can be compared with its source) and used by the
tests to pin `oracle.maps` bit-exactly for even sizes:

* kernel form  src/mapping_cpu_kernels.jl (DFT_to_beta :11-351, beta_to_DFT :363-738)
* loop form    src/mapping_cpu.jl         (DFT_to_beta :2-186,  beta_to_DFT :189-497)
* block form   test/fft_wei.jl            (DFT_to_beta :32-132, beta_to_DFT :145-278)

Pure Python loops: meant for the small shapes of the reference's unit tests.
"""
from __future__ import annotations

import numpy as np

S2 = float(np.sqrt(np.float64(2.0)))


def _c(re, im):
    return complex(float(re), float(im))


def _div(re, im):
    """Julia `(re + im*b) / sqrt(2)` evaluated per component."""
    return complex(re / S2, im / S2)


# --------------------------------------------------------------------------
# kernel (gather) form: src/mapping_cpu_kernels.jl
# --------------------------------------------------------------------------

def kernel_dft_to_beta(v, size):
    """mapping_cpu_kernels.jl:2-351.  v: numpy array with Julia shape
    (half or full spectrum), indexed 1-based through V()."""
    dim = len(size)
    N = int(np.prod(size))
    beta = np.empty(N)
    if dim == 1:
        Nn = size[0]
        M = Nn // 2

        def V(i):
            return v[i - 1]
        for i in range(1, N + 1):          # :11-21
            if i == 1:
                beta[i - 1] = V(1).real
            elif i == 2:
                beta[i - 1] = V(M + 1).real
            elif 3 <= i <= M + 1:
                beta[i - 1] = S2 * V(i - 1).real
            else:
                beta[i - 1] = S2 * V(i - M).imag
        return beta
    if dim == 2:
        N1, N2 = size
        M1, M2 = N1 // 2, N2 // 2
        P1, P2 = M1 - 1, M2 - 1
        PP = P1 * P2

        def V(a, b):
            return v[a - 1, b - 1]
        for i in range(1, N + 1):          # :38-87
            if i == 1:
                x = V(1, 1).real
            elif i == 2:
                x = V(1, M2 + 1).real
            elif i == 3:
                x = V(M1 + 1, 1).real
            elif i == 4:
                x = V(M1 + 1, M2 + 1).real
            elif 5 <= i <= 4 + P2:
                x = S2 * V(1, i - 3).real
            elif 5 + P2 <= i <= 4 + 2 * P2:
                x = S2 * V(1, i - P2 - 3).imag
            elif 5 + 2 * P2 <= i <= 4 + 3 * P2:
                x = S2 * V(M1 + 1, i - 2 * P2 - 3).real
            elif 5 + 3 * P2 <= i <= 4 + 4 * P2:
                x = S2 * V(M1 + 1, i - 3 * P2 - 3).imag
            elif 5 + 4 * P2 <= i <= 4 + 4 * P2 + P1:
                x = S2 * V(i - 4 * P2 - 3, 1).real
            elif 5 + 4 * P2 + P1 <= i <= 4 + 4 * P2 + 2 * P1:
                x = S2 * V(i - 4 * P2 - P1 - 3, 1).imag
            elif 5 + 4 * P2 + 2 * P1 <= i <= 4 + 4 * P2 + 3 * P1:
                x = S2 * V(i - 4 * P2 - 2 * P1 - 3, M2 + 1).real
            elif 5 + 4 * P2 + 3 * P1 <= i <= 4 + 4 * P2 + 4 * P1:
                x = S2 * V(i - 4 * P2 - 3 * P1 - 3, M2 + 1).imag
            else:
                base = 4 + 4 * P2 + 4 * P1
                blk, j = divmod(i - base - 1, PP)
                j += 1
                i2 = (j - 1) // P1 + 1
                i1 = (j - 1) % P1 + 1
                if blk == 0:
                    x = S2 * V(i1 + 1, i2 + 1).real
                elif blk == 1:
                    x = S2 * V(i1 + 1, i2 + 1).imag
                elif blk == 2:
                    x = S2 * V(i1 + 1, i2 + (M2 + 1)).real
                else:
                    x = S2 * V(i1 + 1, i2 + (M2 + 1)).imag
            beta[i - 1] = x
        return beta
    # dim == 3, :109-351
    N1, N2, N3 = size
    M1, M2, M3 = N1 // 2, N2 // 2, N3 // 2
    P1, P2, P3 = M1 - 1, M2 - 1, M3 - 1
    P23, P13, P12, P123 = P2 * P3, P1 * P3, P1 * P2, P1 * P2 * P3

    def V(a, b, c):
        return v[a - 1, b - 1, c - 1]
    verts = [(1, 1, 1), (1, 1, M3 + 1), (1, M2 + 1, 1), (1, M2 + 1, M3 + 1),
             (M1 + 1, 1, 1), (M1 + 1, 1, M3 + 1), (M1 + 1, M2 + 1, 1),
             (M1 + 1, M2 + 1, M3 + 1)]
    ek = [(1, 1), (1, M2 + 1), (M1 + 1, 1), (M1 + 1, M2 + 1)]
    ej = [(1, 1), (1, M3 + 1), (M1 + 1, 1), (M1 + 1, M3 + 1)]
    ei = [(1, 1), (1, M3 + 1), (M2 + 1, 1), (M2 + 1, M3 + 1)]
    for i in range(1, N + 1):
        if i <= 8:                                        # :110-125
            beta[i - 1] = V(*verts[i - 1]).real
            continue
        j = i - 8
        if j <= 8 * P3:                                   # :126-141
            g, t = divmod(j - 1, P3)
            a, b = ek[g // 2]
            z = V(a, b, t + 2)
            beta[i - 1] = S2 * (z.real if g % 2 == 0 else z.imag)
            continue
        j -= 8 * P3
        if j <= 8 * P2:                                   # :142-157
            g, t = divmod(j - 1, P2)
            a, c = ej[g // 2]
            z = V(a, t + 2, c)
            beta[i - 1] = S2 * (z.real if g % 2 == 0 else z.imag)
            continue
        j -= 8 * P2
        if j <= 8 * P1:                                   # :158-173
            g, t = divmod(j - 1, P1)
            b, c = ei[g // 2]
            z = V(t + 2, b, c)
            beta[i - 1] = S2 * (z.real if g % 2 == 0 else z.imag)
            continue
        j -= 8 * P1
        if j <= 8 * P23:                                  # :174-213
            g, t = divmod(j - 1, P23)
            i3 = t // P2 + 1
            i2 = t % P2 + 1
            row = 1 if g < 4 else M1 + 1
            col = i2 + 1 if g % 4 < 2 else i2 + (M2 + 1)
            z = V(row, col, i3 + 1)
            beta[i - 1] = S2 * (z.real if g % 2 == 0 else z.imag)
            continue
        j -= 8 * P23
        if j <= 8 * P13:                                  # :214-253
            g, t = divmod(j - 1, P13)
            i3 = t // P1 + 1
            i1 = t % P1 + 1
            col = 1 if g < 4 else M2 + 1
            if g % 4 == 0:
                beta[i - 1] = S2 * V(i1 + 1, col, i3 + 1).real
            elif g % 4 == 1:
                beta[i - 1] = S2 * V(i1 + 1, col, i3 + 1).imag
            elif g % 4 == 2:
                beta[i - 1] = S2 * V(M1 + 1 - i1, col, N3 + 1 - i3).real
            else:
                beta[i - 1] = -S2 * V(M1 + 1 - i1, col, N3 + 1 - i3).imag
            continue
        j -= 8 * P13
        if j <= 8 * P12:                                  # :254-293
            g, t = divmod(j - 1, P12)
            i2 = t // P1 + 1
            i1 = t % P1 + 1
            dep = 1 if g < 4 else M3 + 1
            if g % 4 == 0:
                beta[i - 1] = S2 * V(i1 + 1, i2 + 1, dep).real
            elif g % 4 == 1:
                beta[i - 1] = S2 * V(i1 + 1, i2 + 1, dep).imag
            elif g % 4 == 2:
                beta[i - 1] = S2 * V(M1 + 1 - i1, N2 + 1 - i2, dep).real
            else:
                beta[i - 1] = -S2 * V(M1 + 1 - i1, N2 + 1 - i2, dep).imag
            continue
        j -= 8 * P12                                      # :294-350
        g, t = divmod(j - 1, P123)
        i3 = t // P12 + 1
        k = t % P12 + 1
        i2 = (k - 1) // P1 + 1
        i1 = (k - 1) % P1 + 1
        if g == 0:
            x = S2 * V(i1 + 1, i2 + 1, i3 + 1).real
        elif g == 1:
            x = S2 * V(i1 + 1, i2 + 1, i3 + 1).imag
        elif g == 2:
            x = S2 * V(M1 + 1 - i1, N2 + 1 - i2, N3 + 1 - i3).real
        elif g == 3:
            x = -S2 * V(M1 + 1 - i1, N2 + 1 - i2, N3 + 1 - i3).imag
        elif g == 4:
            x = S2 * V(i1 + 1, i2 + (M2 + 1), i3 + 1).real
        elif g == 5:
            x = S2 * V(i1 + 1, i2 + (M2 + 1), i3 + 1).imag
        elif g == 6:
            x = S2 * V(M1 + 1 - i1, M2 + 1 - i2, N3 + 1 - i3).real
        else:
            x = -S2 * V(M1 + 1 - i1, M2 + 1 - i2, N3 + 1 - i3).imag
        beta[i - 1] = x
    return beta


def kernel_beta_to_dft(beta, size, rdft=True):
    """mapping_cpu_kernels.jl:353-738: every output element computed by its own
    formula.  Returns a complex array with the Julia shape of the output."""
    b = np.asarray(beta, dtype=np.float64)

    def B(i):
        return b[i - 1]
    dim = len(size)
    if dim == 1:
        N = size[0]
        M = N // 2
        v = np.zeros(M + 1 if rdft else N, np.complex128)
        for i in range(1, (M + 1 if rdft else N) + 1):     # :363-373
            if i == 1:
                v[i - 1] = _c(B(1), 0.0)
            elif i == M + 1:
                v[i - 1] = _c(B(2), 0.0)
            elif 2 <= i <= M:
                v[i - 1] = _div(B(i + 1), B(M + i))
            else:
                v[i - 1] = _div(B(N + 3 - i), -B(N + M + 2 - i))
        return v
    if dim == 2:
        N1, N2 = size
        M1, M2 = N1 // 2, N2 // 2
        P1, P2 = M1 - 1, M2 - 1
        PP = P1 * P2
        rows = M1 + 1 if rdft else N1
        v = np.full((rows, N2), np.nan + 0j)
        for i in range(1, rows + 1):                        # :392-467
            for j in range(1, N2 + 1):
                out = None
                if i == 1:
                    if j == 1:
                        out = _c(B(1), 0)
                    elif j == M2 + 1:
                        out = _c(B(2), 0)
                elif i == M1 + 1:
                    if j == 1:
                        out = _c(B(3), 0)
                    elif j == M2 + 1:
                        out = _c(B(4), 0)
                if i == 1:
                    if 2 <= j <= M2:
                        o = 3
                        out = _div(B(j + o), B(j + o + P2))
                    elif M2 + 2 <= j <= N2:
                        o = N2 + 5
                        out = _div(B(o - j), -B(o + P2 - j))
                elif i == M1 + 1:
                    if 2 <= j <= M2:
                        o = 3 + 2 * P2
                        out = _div(B(j + o), B(j + o + P2))
                    elif M2 + 2 <= j <= N2:
                        o = N2 + 5 + 2 * P2
                        out = _div(B(o - j), -B(o + P2 - j))
                if j == 1:
                    if 2 <= i <= M1:
                        o = 3 + 4 * P2
                        out = _div(B(i + o), B(i + o + P1))
                    elif M1 + 2 <= i <= N1:
                        o = N1 + 5 + 4 * P2
                        out = _div(B(o - i), -B(o + P1 - i))
                elif j == M2 + 1:
                    if 2 <= i <= M1:
                        o = 3 + 4 * P2 + 2 * P1
                        out = _div(B(i + o), B(i + o + P1))
                    elif M1 + 2 <= i <= N1:
                        o = N1 + 5 + 4 * P2 + 2 * P1
                        out = _div(B(o - i), -B(o + P1 - i))
                if 2 <= i <= M1:
                    if 2 <= j <= M2:
                        o = 4 + 4 * P2 + 4 * P1
                        idx = (j - 2) * P1 + (i - 1)
                        out = _div(B(o + idx), B(o + PP + idx))
                    elif M2 + 2 <= j <= N2:
                        o = 4 + 4 * P2 + 4 * P1 + 2 * PP
                        idx = (j - M2 - 2) * P1 + (i - 1)
                        out = _div(B(o + idx), B(o + PP + idx))
                elif M1 + 2 <= i <= N1:
                    if 2 <= j <= M2:
                        o = 4 + 4 * P2 + 4 * P1 + 3 * PP
                        idx = (j - 2) * P1 + (i - M1 - 2)
                        out = _div(B(o - idx), -B(o + PP - idx))
                    elif M2 + 2 <= j <= N2:
                        o = 4 + 4 * P2 + 4 * P1 + PP
                        idx = (j - M2 - 2) * P1 + (i - M1 - 2)
                        out = _div(B(o - idx), -B(o + PP - idx))
                v[i - 1, j - 1] = out
        return v
    # dim 3, :494-738
    N1, N2, N3 = size
    M1, M2, M3 = N1 // 2, N2 // 2, N3 // 2
    P1, P2, P3 = M1 - 1, M2 - 1, M3 - 1
    P23, P13, P12, P123 = P2 * P3, P1 * P3, P1 * P2, P1 * P2 * P3
    rows = M1 + 1 if rdft else N1
    v = np.full((rows, N2, N3), np.nan + 0j)
    F = 8 + 8 * P3 + 8 * P2 + 8 * P1      # start of face blocks
    FJ = F + 8 * P23
    FK = FJ + 8 * P13
    C = FK + 8 * P12
    for i in range(1, rows + 1):
        for j in range(1, N2 + 1):
            for k in range(1, N3 + 1):
                out = None
                # vertices :496-524
                vi = {1: 0, M1 + 1: 4}.get(i)
                vj = {1: 0, M2 + 1: 2}.get(j)
                vk = {1: 0, M3 + 1: 1}.get(k)
                if vi is not None and vj is not None and vk is not None:
                    out = _c(B(1 + vi + vj + vk), 0)
                # edges along k :527-555
                if i in (1, M1 + 1) and j in (1, M2 + 1):
                    g = (0 if i == 1 else 4) + (0 if j == 1 else 2)
                    if 2 <= k <= M3:
                        out = _div(B(k + 7 + g * P3), B(k + 7 + (g + 1) * P3))
                    elif M3 + 2 <= k <= N3:
                        out = _div(B(N3 + 9 + g * P3 - k), -B(N3 + 9 + (g + 1) * P3 - k))
                # edges along j :557-585
                if i in (1, M1 + 1) and k in (1, M3 + 1):
                    g = (0 if i == 1 else 4) + (0 if k == 1 else 2)
                    o = 8 * P3
                    if 2 <= j <= M2:
                        out = _div(B(j + 7 + o + g * P2), B(j + 7 + o + (g + 1) * P2))
                    elif M2 + 2 <= j <= N2:
                        out = _div(B(N2 + 9 + o + g * P2 - j), -B(N2 + 9 + o + (g + 1) * P2 - j))
                # edges along i :587-615
                if j in (1, M2 + 1) and k in (1, M3 + 1):
                    g = (0 if j == 1 else 4) + (0 if k == 1 else 2)
                    o = 8 * P3 + 8 * P2
                    if 2 <= i <= M1:
                        out = _div(B(i + 7 + o + g * P1), B(i + 7 + o + (g + 1) * P1))
                    elif M1 + 2 <= i <= N1:
                        out = _div(B(N1 + 9 + o + g * P1 - i), -B(N1 + 9 + o + (g + 1) * P1 - i))
                # faces i in {1, M1+1} :618-646
                if i in (1, M1 + 1):
                    g = 0 if i == 1 else 4
                    if 2 <= k <= M3:
                        if 2 <= j <= M2:
                            t = (k - 2) * P2 + (j - 1)
                            out = _div(B(F + g * P23 + t), B(F + (g + 1) * P23 + t))
                        elif M2 + 2 <= j <= N2:
                            t = (k - 2) * P2 + (j - (M2 + 1))
                            out = _div(B(F + (g + 2) * P23 + t), B(F + (g + 3) * P23 + t))
                    elif M3 + 2 <= k <= N3:
                        if 2 <= j <= M2:
                            t = (N3 - k) * P2 + (M2 + 1 - j)
                            out = _div(B(F + (g + 2) * P23 + t), -B(F + (g + 3) * P23 + t))
                        elif M2 + 2 <= j <= N2:
                            t = (N3 - k) * P2 + (N2 + 1 - j)
                            out = _div(B(F + g * P23 + t), -B(F + (g + 1) * P23 + t))
                # faces j in {1, M2+1} :648-676
                if j in (1, M2 + 1):
                    g = 0 if j == 1 else 4
                    if 2 <= k <= M3:
                        if 2 <= i <= M1:
                            t = (k - 2) * P1 + (i - 1)
                            out = _div(B(FJ + g * P13 + t), B(FJ + (g + 1) * P13 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (k - 2) * P1 + (i - (M1 + 1))
                            out = _div(B(FJ + (g + 2) * P13 + t), B(FJ + (g + 3) * P13 + t))
                    elif M3 + 2 <= k <= N3:
                        if 2 <= i <= M1:
                            t = (N3 - k) * P1 + (M1 + 1 - i)
                            out = _div(B(FJ + (g + 2) * P13 + t), -B(FJ + (g + 3) * P13 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (N3 - k) * P1 + (N1 + 1 - i)
                            out = _div(B(FJ + g * P13 + t), -B(FJ + (g + 1) * P13 + t))
                # faces k in {1, M3+1} :678-706
                if k in (1, M3 + 1):
                    g = 0 if k == 1 else 4
                    if 2 <= j <= M2:
                        if 2 <= i <= M1:
                            t = (j - 2) * P1 + (i - 1)
                            out = _div(B(FK + g * P12 + t), B(FK + (g + 1) * P12 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (j - 2) * P1 + (i - (M1 + 1))
                            out = _div(B(FK + (g + 2) * P12 + t), B(FK + (g + 3) * P12 + t))
                    elif M2 + 2 <= j <= N2:
                        if 2 <= i <= M1:
                            t = (N2 - j) * P1 + (M1 + 1 - i)
                            out = _div(B(FK + (g + 2) * P12 + t), -B(FK + (g + 3) * P12 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (N2 - j) * P1 + (N1 + 1 - i)
                            out = _div(B(FK + g * P12 + t), -B(FK + (g + 1) * P12 + t))
                # interior :709-737
                if 2 <= k <= M3:
                    if 2 <= j <= M2:
                        if 2 <= i <= M1:
                            t = (k - 2) * P12 + (j - 2) * P1 + (i - 1)
                            out = _div(B(C + t), B(C + P123 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (k - 2) * P12 + (j - 2) * P1 + (i - (M1 + 1))
                            out = _div(B(C + 2 * P123 + t), B(C + 3 * P123 + t))
                    elif M2 + 2 <= j <= N2:
                        if 2 <= i <= M1:
                            t = (k - 2) * P12 + (j - (M2 + 2)) * P1 + (i - 1)
                            out = _div(B(C + 4 * P123 + t), B(C + 5 * P123 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (k - 2) * P12 + (j - (M2 + 2)) * P1 + (i - (M1 + 1))
                            out = _div(B(C + 6 * P123 + t), B(C + 7 * P123 + t))
                elif M3 + 2 <= k <= N3:
                    if 2 <= j <= M2:
                        if 2 <= i <= M1:
                            t = (N3 - k) * P12 + (M2 - j) * P1 + (M1 + 1 - i)
                            out = _div(B(C + 6 * P123 + t), -B(C + 7 * P123 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (N3 - k) * P12 + (M2 - j) * P1 + (N1 + 1 - i)
                            out = _div(B(C + 4 * P123 + t), -B(C + 5 * P123 + t))
                    elif M2 + 2 <= j <= N2:
                        if 2 <= i <= M1:
                            t = (N3 - k) * P12 + (N2 - j) * P1 + (M1 + 1 - i)
                            out = _div(B(C + 2 * P123 + t), -B(C + 3 * P123 + t))
                        elif M1 + 2 <= i <= N1:
                            t = (N3 - k) * P12 + (N2 - j) * P1 + (N1 + 1 - i)
                            out = _div(B(C + t), -B(C + P123 + t))
                v[i - 1, j - 1, k - 1] = out
    return v


# --------------------------------------------------------------------------
# loop form: src/mapping_cpu.jl
# --------------------------------------------------------------------------

def loop_dft_to_beta(v, size):
    """mapping_cpu.jl:2-186."""
    dim = len(size)
    N = int(np.prod(size))
    beta = np.empty(N)

    def put(l, x):
        beta[l - 1] = x
    if dim == 1:
        M = size[0] // 2
        put(1, v[0].real)
        put(2, v[M].real)
        for i in range(2, M + 1):
            put(i + 1, S2 * v[i - 1].real)
            put(i + M, S2 * v[i - 1].imag)
        return beta
    if dim == 2:
        N1, N2 = size
        M1, M2 = N1 // 2, N2 // 2
        P1, P2 = M1 - 1, M2 - 1
        PP = P1 * P2
        V = lambda a, b: v[a - 1, b - 1]          # noqa: E731
        put(1, V(1, 1).real); put(2, V(1, M2 + 1).real)
        put(3, V(M1 + 1, 1).real); put(4, V(M1 + 1, M2 + 1).real)
        index = 4
        for l in range(index + 1, index + P2 + 1):
            t = l - index + 1
            put(l, S2 * V(1, t).real); put(l + P2, S2 * V(1, t).imag)
            put(l + 2 * P2, S2 * V(M1 + 1, t).real); put(l + 3 * P2, S2 * V(M1 + 1, t).imag)
        index += 4 * P2
        for l in range(index + 1, index + P1 + 1):
            t = l - index + 1
            put(l, S2 * V(t, 1).real); put(l + P1, S2 * V(t, 1).imag)
            put(l + 2 * P1, S2 * V(t, M2 + 1).real); put(l + 3 * P1, S2 * V(t, M2 + 1).imag)
        index += 4 * P1
        l = index
        for j in range(2, M2 + 1):
            for i in range(2, M1 + 1):
                l += 1
                put(l, S2 * V(i, j).real); put(l + PP, S2 * V(i, j).imag)
                put(l + 2 * PP, S2 * V(i, j + M2).real); put(l + 3 * PP, S2 * V(i, j + M2).imag)
        return beta
    N1, N2, N3 = size
    M1, M2, M3 = N1 // 2, N2 // 2, N3 // 2
    P1, P2, P3 = M1 - 1, M2 - 1, M3 - 1
    P23, P13, P12, P123 = P2 * P3, P1 * P3, P1 * P2, P1 * P2 * P3
    V = lambda a, b, c: v[a - 1, b - 1, c - 1]    # noqa: E731
    for n, (a, b, c) in enumerate([(1, 1, 1), (1, 1, M3 + 1), (1, M2 + 1, 1), (1, M2 + 1, M3 + 1),
                                   (M1 + 1, 1, 1), (M1 + 1, 1, M3 + 1), (M1 + 1, M2 + 1, 1),
                                   (M1 + 1, M2 + 1, M3 + 1)]):
        put(n + 1, V(a, b, c).real)
    index = 8
    for l in range(index + 1, index + P3 + 1):
        t = l - index + 1
        for q, (a, b) in enumerate([(1, 1), (1, M2 + 1), (M1 + 1, 1), (M1 + 1, M2 + 1)]):
            put(l + 2 * q * P3, S2 * V(a, b, t).real)
            put(l + (2 * q + 1) * P3, S2 * V(a, b, t).imag)
    index += 8 * P3
    for l in range(index + 1, index + P2 + 1):
        t = l - index + 1
        for q, (a, c) in enumerate([(1, 1), (1, M3 + 1), (M1 + 1, 1), (M1 + 1, M3 + 1)]):
            put(l + 2 * q * P2, S2 * V(a, t, c).real)
            put(l + (2 * q + 1) * P2, S2 * V(a, t, c).imag)
    index += 8 * P2
    for l in range(index + 1, index + P1 + 1):
        t = l - index + 1
        for q, (b, c) in enumerate([(1, 1), (1, M3 + 1), (M2 + 1, 1), (M2 + 1, M3 + 1)]):
            put(l + 2 * q * P1, S2 * V(t, b, c).real)
            put(l + (2 * q + 1) * P1, S2 * V(t, b, c).imag)
    index += 8 * P1
    l = index
    for k in range(2, M3 + 1):
        for j in range(2, M2 + 1):
            l += 1
            put(l, S2 * V(1, j, k).real); put(l + P23, S2 * V(1, j, k).imag)
            put(l + 2 * P23, S2 * V(1, j + M2, k).real); put(l + 3 * P23, S2 * V(1, j + M2, k).imag)
            put(l + 4 * P23, S2 * V(M1 + 1, j, k).real); put(l + 5 * P23, S2 * V(M1 + 1, j, k).imag)
            put(l + 6 * P23, S2 * V(M1 + 1, j + M2, k).real)
            put(l + 7 * P23, S2 * V(M1 + 1, j + M2, k).imag)
    index += 8 * P23
    l = index
    for k in range(2, M3 + 1):
        for i in range(2, M1 + 1):
            l += 1
            put(l, S2 * V(i, 1, k).real); put(l + P13, S2 * V(i, 1, k).imag)
            put(l + 2 * P13, S2 * V(N1 - i - M1 + 2, 1, N3 - k + 2).real)
            put(l + 3 * P13, -S2 * V(N1 - i - M1 + 2, 1, N3 - k + 2).imag)
            put(l + 4 * P13, S2 * V(i, M2 + 1, k).real); put(l + 5 * P13, S2 * V(i, M2 + 1, k).imag)
            put(l + 6 * P13, S2 * V(N1 - i - M1 + 2, M2 + 1, N3 - k + 2).real)
            put(l + 7 * P13, -S2 * V(N1 - i - M1 + 2, M2 + 1, N3 - k + 2).imag)
    index += 8 * P13
    l = index
    for j in range(2, M2 + 1):
        for i in range(2, M1 + 1):
            l += 1
            put(l, S2 * V(i, j, 1).real); put(l + P12, S2 * V(i, j, 1).imag)
            put(l + 2 * P12, S2 * V(N1 - i - M1 + 2, N2 - j + 2, 1).real)
            put(l + 3 * P12, -S2 * V(N1 - i - M1 + 2, N2 - j + 2, 1).imag)
            put(l + 4 * P12, S2 * V(i, j, M3 + 1).real); put(l + 5 * P12, S2 * V(i, j, M3 + 1).imag)
            put(l + 6 * P12, S2 * V(N1 - i - M1 + 2, N2 - j + 2, M3 + 1).real)
            put(l + 7 * P12, -S2 * V(N1 - i - M1 + 2, N2 - j + 2, M3 + 1).imag)
    index += 8 * P12
    l = index
    for k in range(2, M3 + 1):
        for j in range(2, M2 + 1):
            for i in range(2, M1 + 1):
                l += 1
                put(l, S2 * V(i, j, k).real); put(l + P123, S2 * V(i, j, k).imag)
                put(l + 2 * P123, S2 * V(N1 - i - M1 + 2, N2 - j + 2, N3 - k + 2).real)
                put(l + 3 * P123, -S2 * V(N1 - i - M1 + 2, N2 - j + 2, N3 - k + 2).imag)
                put(l + 4 * P123, S2 * V(i, j + M2, k).real)
                put(l + 5 * P123, S2 * V(i, j + M2, k).imag)
                put(l + 6 * P123, S2 * V(N1 - i - M1 + 2, N2 - j - M2 + 2, N3 - k + 2).real)
                put(l + 7 * P123, -S2 * V(N1 - i - M1 + 2, N2 - j - M2 + 2, N3 - k + 2).imag)
    return beta


def loop_beta_to_dft(beta, size, rdft=True):
    """mapping_cpu.jl:189-497 (conjugate fills copy already-written entries)."""
    b = np.asarray(beta, dtype=np.float64)
    B = lambda i: b[i - 1]                          # noqa: E731
    dim = len(size)
    if dim == 1:
        N = size[0]
        M = N // 2
        v = np.full(M + 1 if rdft else N, np.nan + 0j)
        v[0] = _c(B(1), 0)
        v[M] = _c(B(2), 0)
        for i in range(2, M + 1):
            v[i - 1] = _div(B(i + 1), B(M + i))
        if not rdft:
            for i in range(2, M + 1):
                v[N - i + 1] = _div(B(i + 1), -B(M + i))
        return v
    if dim == 2:
        N1, N2 = size
        M1, M2 = N1 // 2, N2 // 2
        P1, P2 = M1 - 1, M2 - 1
        v = np.full((M1 + 1 if rdft else N1, N2), np.nan + 0j)

        def S(a, c, x):
            v[a - 1, c - 1] = x
        G = lambda a, c: v[a - 1, c - 1]            # noqa: E731
        S(1, 1, _c(B(1), 0)); S(1, M2 + 1, _c(B(2), 0))
        S(M1 + 1, 1, _c(B(3), 0)); S(M1 + 1, M2 + 1, _c(B(4), 0))
        for i in range(1, P1 + 1):
            S(i + 1, 1, _div(B(4 + 4 * P2 + i), B(4 + 4 * P2 + P1 + i)))
            S(i + 1, M2 + 1, _div(B(4 + 4 * P2 + 2 * P1 + i), B(4 + 4 * P2 + 3 * P1 + i)))
        for i in range(1, P2 + 1):
            S(1, i + 1, _div(B(4 + i), B(4 + P2 + i)))
            S(M1 + 1, i + 1, _div(B(4 + 2 * P2 + i), B(4 + 3 * P2 + i)))
        j = 0
        c = 0
        for col in range(2, M2 + 1):
            for row in range(2, M1 + 1):
                j += 1
                S(row, col, _div(B(4 + 4 * P2 + 4 * P1 + j), B(4 + 4 * P2 + 4 * P1 + P1 * P2 + j)))
                S(M1 - row + 2, N2 - col + 2,
                  _div(B(4 + 4 * P2 + 4 * P1 + 3 * P1 * P2 - c), B(N1 * N2 - c)))
                c += 1
        for i in range(1, P2 + 1):
            S(1, M2 + 1 + i, np.conj(G(1, M2 - i + 1)))
            S(M1 + 1, M2 + 1 + i, np.conj(G(M1 + 1, M2 - i + 1)))
        if not rdft:
            for i in range(1, P1 + 1):
                S(N1 - i + 1, 1, _div(B(4 + 4 * P2 + i), -B(4 + 4 * P2 + P1 + i)))
                S(N1 - i + 1, M2 + 1, _div(B(4 + 4 * P2 + 2 * P1 + i), -B(4 + 4 * P2 + 3 * P1 + i)))
                for jj in range(1, P2 + 1):
                    S(N1 - i + 1, M2 - jj + 1, np.conj(G(i + 1, M2 + jj + 1)))
                    S(M1 + i + 1, M2 + jj + 1, np.conj(G(M1 - i + 1, M2 - jj + 1)))
        return v
    N1, N2, N3 = size
    M1, M2, M3 = N1 // 2, N2 // 2, N3 // 2
    P1, P2, P3 = M1 - 1, M2 - 1, M3 - 1
    v = np.full((M1 + 1 if rdft else N1, N2, N3), np.nan + 0j)

    def S(a, c, d, x):
        v[a - 1, c - 1, d - 1] = x
    G = lambda a, c, d: v[a - 1, c - 1, d - 1]       # noqa: E731
    for n, (a, c, d) in enumerate([(1, 1, 1), (1, 1, M3 + 1), (1, M2 + 1, 1), (1, M2 + 1, M3 + 1),
                                   (M1 + 1, 1, 1), (M1 + 1, 1, M3 + 1), (M1 + 1, M2 + 1, 1),
                                   (M1 + 1, M2 + 1, M3 + 1)]):
        S(a, c, d, _c(B(n + 1), 0))
    o = 8 + 8 * P3 + 8 * P2
    for i in range(1, P1 + 1):
        for q, (c, d) in enumerate([(1, 1), (1, M3 + 1), (M2 + 1, 1), (M2 + 1, M3 + 1)]):
            S(i + 1, c, d, _div(B(o + 2 * q * P1 + i), B(o + (2 * q + 1) * P1 + i)))
    o = 8 + 8 * P3
    for j in range(1, P2 + 1):
        for q, (a, d) in enumerate([(1, 1), (1, M3 + 1), (M1 + 1, 1), (M1 + 1, M3 + 1)]):
            S(a, j + 1, d, _div(B(o + 2 * q * P2 + j), B(o + (2 * q + 1) * P2 + j)))
    for k in range(1, P3 + 1):
        for q, (a, c) in enumerate([(1, 1), (1, M2 + 1), (M1 + 1, 1), (M1 + 1, M2 + 1)]):
            S(a, c, k + 1, _div(B(8 + 2 * q * P3 + k), B(8 + (2 * q + 1) * P3 + k)))
    o = 8 + 8 * P3 + 8 * P2 + 8 * P1 + 8 * P2 * P3 + 8 * P1 * P3
    for j in range(1, P2 + 1):
        for i in range(1, P1 + 1):
            ij = (j - 1) * P1 + i
            S(i + 1, j + 1, 1, _div(B(o + ij), B(o + P1 * P2 + ij)))
            S(M1 - i + 1, N2 - j + 1, 1, _div(B(o + 2 * P1 * P2 + ij), -B(o + 3 * P1 * P2 + ij)))
            S(i + 1, j + 1, M3 + 1, _div(B(o + 4 * P1 * P2 + ij), B(o + 5 * P1 * P2 + ij)))
            S(M1 - i + 1, N2 - j + 1, M3 + 1, _div(B(o + 6 * P1 * P2 + ij), -B(o + 7 * P1 * P2 + ij)))
    o = 8 + 8 * P3 + 8 * P2 + 8 * P1 + 8 * P2 * P3
    for k in range(1, P3 + 1):
        for i in range(1, P1 + 1):
            ik = (k - 1) * P1 + i
            S(i + 1, 1, k + 1, _div(B(o + ik), B(o + P1 * P3 + ik)))
            S(M1 - i + 1, 1, N3 - k + 1, _div(B(o + 2 * P1 * P3 + ik), -B(o + 3 * P1 * P3 + ik)))
            S(i + 1, M2 + 1, k + 1, _div(B(o + 4 * P1 * P3 + ik), B(o + 5 * P1 * P3 + ik)))
            S(M1 - i + 1, M2 + 1, N3 - k + 1, _div(B(o + 6 * P1 * P3 + ik), -B(o + 7 * P1 * P3 + ik)))
    o = 8 + 8 * P3 + 8 * P2 + 8 * P1
    for k in range(1, P3 + 1):
        for j in range(1, P2 + 1):
            jk = (k - 1) * P2 + j
            S(1, j + 1, k + 1, _div(B(o + jk), B(o + P2 * P3 + jk)))
            S(1, M2 + j + 1, k + 1, _div(B(o + 2 * P2 * P3 + jk), B(o + 3 * P2 * P3 + jk)))
            S(M1 + 1, j + 1, k + 1, _div(B(o + 4 * P2 * P3 + jk), B(o + 5 * P2 * P3 + jk)))
            S(M1 + 1, M2 + j + 1, k + 1, _div(B(o + 6 * P2 * P3 + jk), B(o + 7 * P2 * P3 + jk)))
    for j in range(1, P2 + 1):
        for a in (1, M1 + 1):
            for d in (1, M3 + 1):
                S(a, M2 + j + 1, d, np.conj(G(a, M2 - j + 1, d)))
    for k in range(1, P3 + 1):
        for a in (1, M1 + 1):
            for c in (1, M2 + 1):
                S(a, c, M3 + 1 + k, np.conj(G(a, c, M3 - k + 1)))
    for k in range(1, P3 + 1):
        for j in range(1, P2 + 1):
            for a in (1, M1 + 1):
                S(a, j + 1, M3 + k + 1, np.conj(G(a, N2 - j + 1, M3 - k + 1)))
                S(a, M2 + j + 1, M3 + k + 1, np.conj(G(a, M2 - j + 1, M3 - k + 1)))
    o = 8 + 8 * P3 + 8 * P2 + 8 * P1 + 8 * P2 * P3 + 8 * P1 * P3 + 8 * P1 * P2
    P123 = P1 * P2 * P3
    for k in range(1, P3 + 1):
        for j in range(1, P2 + 1):
            for i in range(1, P1 + 1):
                ijk = (k - 1) * P2 * P1 + (j - 1) * P1 + i
                S(i + 1, j + 1, k + 1, _div(B(o + ijk), B(o + P123 + ijk)))
                S(M1 - i + 1, N2 - j + 1, N3 - k + 1, _div(B(o + 2 * P123 + ijk), -B(o + 3 * P123 + ijk)))
                S(i + 1, M2 + j + 1, k + 1, _div(B(o + 4 * P123 + ijk), B(o + 5 * P123 + ijk)))
                S(M1 - i + 1, M2 - j + 1, N3 - k + 1, _div(B(o + 6 * P123 + ijk), -B(o + 7 * P123 + ijk)))
    if not rdft:
        o = 8 + 8 * P3 + 8 * P2
        for i in range(1, P1 + 1):
            for q, (c, d) in enumerate([(1, 1), (1, M3 + 1), (M2 + 1, 1), (M2 + 1, M3 + 1)]):
                S(M1 + 1 + i, c, d, _div(B(o + (2 * q + 1) * P1 - i + 1), -B(o + (2 * q + 2) * P1 - i + 1)))
        for j in range(1, P2 + 1):
            for i in range(1, P1 + 1):
                for d in (1, M3 + 1):
                    S(N1 - i + 1, M2 - j + 1, d, np.conj(G(i + 1, M2 + j + 1, d)))
                    S(M1 + i + 1, M2 + j + 1, d, np.conj(G(M1 - i + 1, M2 - j + 1, d)))
        for k in range(1, P3 + 1):
            for i in range(1, P1 + 1):
                S(N1 - i + 1, 1, M3 - k + 1, np.conj(G(i + 1, 1, M3 + 1 + k)))
                S(N1 - i + 1, M2 + 1, M3 - k + 1, np.conj(G(i + 1, M2 + 1, M3 + k + 1)))
                S(M1 + i + 1, 1, M3 + k + 1, np.conj(G(M1 - i + 1, 1, M3 - k + 1)))
                S(M1 + i + 1, M2 + 1, M3 + k + 1, np.conj(G(M1 - i + 1, M2 + 1, M3 - k + 1)))
        for k in range(1, P3 + 1):
            for j in range(1, P2 + 1):
                for i in range(1, P1 + 1):
                    S(M1 + i + 1, M2 + j + 1, M3 + k + 1, np.conj(G(M1 - i + 1, M2 - j + 1, M3 - k + 1)))
                    S(N1 - i + 1, M2 - j + 1, M3 - k + 1, np.conj(G(i + 1, M2 + j + 1, M3 + k + 1)))
                    S(M1 + i + 1, j + 1, M3 + k + 1, np.conj(G(M1 - i + 1, N2 - j + 1, M3 - k + 1)))
                    S(N1 - i + 1, N2 - j + 1, M3 - k + 1, np.conj(G(i + 1, j + 1, M3 + k + 1)))
    return v


# --------------------------------------------------------------------------
# block (concatenation) form: test/fft_wei.jl
# --------------------------------------------------------------------------

def _F(a):
    return np.asarray(a).reshape(-1, order="F")


def wei_dft_to_beta(v, size):
    """fft_wei.jl:32-132; v is the FULL spectrum (Julia shape)."""
    dim = len(size)
    if dim == 1:
        M = size[0] // 2
        return np.concatenate([[v[0].real], [v[M].real], S2 * v[1:M].real, S2 * v[1:M].imag])
    if dim == 2:
        N1, N2 = size
        M1, M2 = N1 // 2, N2 // 2
        parts = [[v[0, 0].real], [v[0, M2].real], [v[M1, 0].real], [v[M1, M2].real],
                 S2 * v[0, 1:M2].real, S2 * v[0, 1:M2].imag,
                 S2 * v[M1, 1:M2].real, S2 * v[M1, 1:M2].imag,
                 S2 * v[1:M1, 0].real, S2 * v[1:M1, 0].imag,
                 S2 * v[1:M1, M2].real, S2 * v[1:M1, M2].imag,
                 S2 * _F(v[1:M1, 1:M2].real), S2 * _F(v[1:M1, 1:M2].imag),
                 S2 * _F(v[1:M1, M2 + 1:N2].real), S2 * _F(v[1:M1, M2 + 1:N2].imag)]
        return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])
    N1, N2, N3 = size
    M1, M2, M3 = N1 // 2, N2 // 2, N3 // 2
    parts = [[v[0, 0, 0].real], [v[0, 0, M3].real], [v[0, M2, 0].real], [v[0, M2, M3].real],
             [v[M1, 0, 0].real], [v[M1, 0, M3].real], [v[M1, M2, 0].real], [v[M1, M2, M3].real]]
    for a, b in [(0, 0), (0, M2), (M1, 0), (M1, M2)]:
        parts += [S2 * v[a, b, 1:M3].real, S2 * v[a, b, 1:M3].imag]
    for a, c in [(0, 0), (0, M3), (M1, 0), (M1, M3)]:
        parts += [S2 * v[a, 1:M2, c].real, S2 * v[a, 1:M2, c].imag]
    for b, c in [(0, 0), (0, M3), (M2, 0), (M2, M3)]:
        parts += [S2 * v[1:M1, b, c].real, S2 * v[1:M1, b, c].imag]
    for a in (0, M1):
        parts += [_F(S2 * v[a, 1:M2, 1:M3].real), _F(S2 * v[a, 1:M2, 1:M3].imag),
                  _F(S2 * v[a, M2 + 1:N2, 1:M3].real), _F(S2 * v[a, M2 + 1:N2, 1:M3].imag)]
    for b in (0, M2):
        parts += [_F(S2 * v[1:M1, b, 1:M3].real), _F(S2 * v[1:M1, b, 1:M3].imag),
                  _F(S2 * v[M1 + 1:N1, b, 1:M3].real), _F(S2 * v[M1 + 1:N1, b, 1:M3].imag)]
    for c in (0, M3):
        parts += [_F(S2 * v[1:M1, 1:M2, c].real), _F(S2 * v[1:M1, 1:M2, c].imag),
                  _F(S2 * v[M1 + 1:N1, 1:M2, c].real), _F(S2 * v[M1 + 1:N1, 1:M2, c].imag)]
    for sa, sb in [(slice(1, M1), slice(1, M2)), (slice(M1 + 1, N1), slice(1, M2)),
                   (slice(1, M1), slice(M2 + 1, N2)), (slice(M1 + 1, N1), slice(M2 + 1, N2))]:
        parts += [_F(S2 * v[sa, sb, 1:M3].real), _F(S2 * v[sa, sb, 1:M3].imag)]
    return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])


def _cv(re, im):
    out = np.empty(np.broadcast(re, im).shape, np.complex128)
    out.real = re / S2
    out.imag = im / S2
    return out


def wei_beta_to_dft(beta, size):
    """fft_wei.jl:145-278: builds the FULL spectrum."""
    b = np.asarray(beta, dtype=np.float64)
    dim = len(size)

    def seg(lo, hi):              # beta[lo:hi] in 1-based inclusive notation
        return b[lo - 1:hi]
    if dim == 1:
        N = size[0]
        M = N // 2
        pos = _cv(seg(3, M + 1), seg(M + 2, N))
        neg = _cv(seg(3, M + 1), -seg(M + 2, N))
        return np.concatenate([[complex(b[0], 0)], pos, [complex(b[1], 0)], neg[::-1]])
    if dim == 2:
        N1, N2 = size
        M1, M2 = N1 // 2, N2 // 2
        P1, P2 = M1 - 1, M2 - 1
        v = np.full((N1, N2), np.nan + 0j)
        c1r, c1i = seg(4 + 4 * P2 + 1, 4 + 4 * P2 + P1), seg(4 + 4 * P2 + P1 + 1, 4 + 4 * P2 + 2 * P1)
        v[:, 0] = np.concatenate([[complex(b[0], 0)], _cv(c1r, c1i), [complex(b[2], 0)],
                                  _cv(c1r, -c1i)[::-1]])
        cen = 4 + 4 * P2 + 4 * P1
        blk = lambda q: seg(cen + q * P1 * P2 + 1, cen + (q + 1) * P1 * P2).reshape(P1, P2, order="F")  # noqa
        v[0, 1:M2] = _cv(seg(5, 4 + P2), seg(4 + P2 + 1, 4 + 2 * P2))
        v[1:M1, 1:M2] = _cv(blk(0), blk(1))
        v[M1, 1:M2] = _cv(seg(4 + 2 * P2 + 1, 4 + 3 * P2), seg(4 + 3 * P2 + 1, 4 + 4 * P2))
        v[M1 + 1:N1, 1:M2] = _cv(blk(2), -blk(3))[::-1, ::-1]
        cmr = seg(4 + 4 * P2 + 2 * P1 + 1, 4 + 4 * P2 + 3 * P1)
        cmi = seg(4 + 4 * P2 + 3 * P1 + 1, 4 + 4 * P2 + 4 * P1)
        v[:, M2] = np.concatenate([[complex(b[1], 0)], _cv(cmr, cmi), [complex(b[3], 0)],
                                   _cv(cmr, -cmi)[::-1]])
        v[0, M2 + 1:N2] = np.conj(v[0, 1:M2])[::-1]
        v[1:M1, M2 + 1:N2] = np.conj(v[M1 + 1:N1, 1:M2])[::-1, ::-1]
        v[M1, M2 + 1:N2] = np.conj(v[M1, 1:M2])[::-1]
        v[M1 + 1:N1, M2 + 1:N2] = np.conj(v[1:M1, 1:M2])[::-1, ::-1]
        return v
    N1, N2, N3 = size
    M1, M2, M3 = N1 // 2, N2 // 2, N3 // 2
    P1, P2, P3 = M1 - 1, M2 - 1, M3 - 1
    v = np.full((N1, N2, N3), np.nan + 0j)
    E3, E2, E1 = 8, 8 + 8 * P3, 8 + 8 * P3 + 8 * P2          # edge block starts (0-based)
    FI = E1 + 8 * P1
    FJ = FI + 8 * P2 * P3
    FK = FJ + 8 * P1 * P3
    IN = FK + 8 * P1 * P2

    def blk(start, q, n, shape=None):
        x = b[start + q * n:start + (q + 1) * n]
        return x if shape is None else x.reshape(shape, order="F")

    def line1(v0, vm, q):
        r, i = blk(E1, q, P1), blk(E1, q + 1, P1)
        return np.concatenate([[complex(v0, 0)], _cv(r, i), [complex(vm, 0)], _cv(r, -i)[::-1]])

    def fill_plane(c):
        """fft_wei.jl:189-228 for the k = c plane (c in {0, M3})."""
        qv = 0 if c == 0 else 1          # vertex offset (beta[1] / beta[2] ...)
        qe = 0 if c == 0 else 2
        v[:, 0, c] = line1(b[0 + qv], b[4 + qv], qe)
        jr = lambda q: blk(E2, q, P2)      # noqa: E731
        fk = lambda q: blk(FK, q, P1 * P2, (P1, P2))  # noqa: E731
        qk = 0 if c == 0 else 4
        v[0, 1:M2, c] = _cv(jr(qe), jr(qe + 1))
        v[1:M1, 1:M2, c] = _cv(fk(qk), fk(qk + 1))
        v[M1, 1:M2, c] = _cv(jr(qe + 4), jr(qe + 5))
        v[M1 + 1:N1, 1:M2, c] = _cv(fk(qk + 2), fk(qk + 3))
        v[:, M2, c] = line1(b[2 + qv], b[6 + qv], qe + 4)
        v[0, M2 + 1:N2, c] = np.conj(v[0, 1:M2, c])[::-1]
        v[M1 + 1:N1, M2 + 1:N2, c] = np.conj(v[1:M1, 1:M2, c])[::-1, ::-1]
        v[M1, M2 + 1:N2, c] = np.conj(v[M1, 1:M2, c])[::-1]
        v[1:M1, M2 + 1:N2, c] = np.conj(v[M1 + 1:N1, 1:M2, c])[::-1, ::-1]

    fill_plane(0)
    fill_plane(M3)
    kr = lambda q: blk(E3, q, P3)                           # noqa: E731
    fj = lambda q: blk(FJ, q, P1 * P3, (P1, P3))            # noqa: E731
    v[0, 0, 1:M3] = _cv(kr(0), kr(1))                      # :231-234
    v[1:M1, 0, 1:M3] = _cv(fj(0), fj(1))
    v[M1, 0, 1:M3] = _cv(kr(4), kr(5))
    v[M1 + 1:N1, 0, 1:M3] = _cv(fj(2), fj(3))
    # :236-239, one row group at a time
    v[0, 0, M3 + 1:N3] = np.conj(v[0, 0, 1:M3])[::-1]
    v[1:M1, 0, M3 + 1:N3] = np.conj(v[M1 + 1:N1, 0, 1:M3])[::-1, ::-1]
    v[M1, 0, M3 + 1:N3] = np.conj(v[M1, 0, 1:M3])[::-1]
    v[M1 + 1:N1, 0, M3 + 1:N3] = np.conj(v[1:M1, 0, 1:M3])[::-1, ::-1]
    v[0, M2, 1:M3] = _cv(kr(2), kr(3))                     # :242-247
    v[1:M1, M2, 1:M3] = _cv(fj(4), fj(5))
    v[M1, M2, 1:M3] = _cv(kr(6), kr(7))
    v[M1 + 1:N1, M2, 1:M3] = _cv(fj(6), fj(7))
    v[0, M2, M3 + 1:N3] = np.conj(v[0, M2, 1:M3])[::-1]     # :250-253
    v[1:M1, M2, M3 + 1:N3] = np.conj(v[M1 + 1:N1, M2, 1:M3])[::-1, ::-1]
    v[M1, M2, M3 + 1:N3] = np.conj(v[M1, M2, 1:M3])[::-1]
    v[M1 + 1:N1, M2, M3 + 1:N3] = np.conj(v[1:M1, M2, 1:M3])[::-1, ::-1]
    fi = lambda q: blk(FI, q, P2 * P3, (P2, P3))            # noqa: E731
    for a, q in ((0, 0), (M1, 4)):                         # :255-263
        v[a, 1:M2, 1:M3] = _cv(fi(q), fi(q + 1))
        v[a, M2 + 1:N2, 1:M3] = _cv(fi(q + 2), fi(q + 3))
        v[a, 1:M2, M3 + 1:N3] = np.conj(v[a, M2 + 1:N2, 1:M3])[::-1, ::-1]
        v[a, M2 + 1:N2, M3 + 1:N3] = np.conj(v[a, 1:M2, 1:M3])[::-1, ::-1]
    it = lambda q: blk(IN, q, P1 * P2 * P3, (P1, P2, P3))   # noqa: E731
    v[1:M1, 1:M2, 1:M3] = _cv(it(0), it(1))                 # :265-275
    v[M1 + 1:N1, M2 + 1:N2, M3 + 1:N3] = np.conj(v[1:M1, 1:M2, 1:M3])[::-1, ::-1, ::-1]
    v[M1 + 1:N1, 1:M2, 1:M3] = _cv(it(2), it(3))
    v[1:M1, M2 + 1:N2, M3 + 1:N3] = np.conj(v[M1 + 1:N1, 1:M2, 1:M3])[::-1, ::-1, ::-1]
    v[1:M1, M2 + 1:N2, 1:M3] = _cv(it(4), it(5))
    v[M1 + 1:N1, 1:M2, M3 + 1:N3] = np.conj(v[1:M1, M2 + 1:N2, 1:M3])[::-1, ::-1, ::-1]
    v[M1 + 1:N1, M2 + 1:N2, 1:M3] = _cv(it(6), it(7))
    v[1:M1, 1:M2, M3 + 1:N3] = np.conj(v[M1 + 1:N1, M2 + 1:N2, 1:M3])[::-1, ::-1, ::-1]
    return v
