"""CPU ORACLE — test infrastructure only (tests/ use it as the checker).

Restatement of the reference's air-gap diagnostics
(onelab_interface/magnetostatics_onelab.hpp:4666-4720 for the 2D grid,
5246-5307 for the 3D grid, 460-519 for the traction on a +y face and the
x-mirror projection), as plain loops over the cells in the reference's
order (k, j, i).  Inputs: node coordinates, per-cell B and H [cells][3]
(cells x fastest), per-cell region / positive-conductor booleans and the
cells' current density z component.
"""
import math

TOL = 1.0e-12


def traction_on_positive_y_face(b, h):
    half_trace = 0.5 * (b[0] * h[0] + b[1] * h[1] + b[2] * h[2])
    return (b[0] * h[1], b[1] * h[1] - half_trace, b[2] * h[1])


def mirror_projection(samples, result):
    """samples: (x, z, measure, traction); replaces result["traction"] when
    the faces pair up at ±x with equal z and measure"""
    if not samples:
        return
    # the reference's comparator orders by z (beyond TOL) then x
    import functools

    def cmp(a, b):
        if abs(a[1] - b[1]) > TOL:
            return -1 if a[1] < b[1] else 1
        return -1 if a[0] < b[0] else (1 if a[0] > b[0] else 0)

    s = sorted(samples, key=functools.cmp_to_key(cmp))
    proj = [0.0, 0.0, 0.0]
    left, right = 0, len(s)
    while left < right:
        right -= 1
        lf, rf = s[left], s[right]
        if left == right:
            if abs(lf[0]) > TOL:
                return
            proj[1] += lf[2] * lf[3][1]
            proj[2] += lf[2] * lf[3][2]
            break
        if abs(lf[1] - rf[1]) > TOL or abs(lf[0] + rf[0]) > TOL or abs(lf[2] - rf[2]) > TOL:
            return
        proj[1] += lf[2] * (lf[3][1] + rf[3][1])
        proj[2] += lf[2] * (lf[3][2] + rf[3][2])
        left += 1
    result["traction"] = proj


def air_gap_diagnostics(coords, cell_b, cell_h, in_region, positive, jz, length_z=None, z_value=0.0):
    """coords: (x, y) for the 2D grid (then length_z is the extrusion) or
    (x, y, z); returns faces, surface measure, flux, traction (3), traction
    magnitude integral, current and conductor cell count"""
    x, y = coords[0], coords[1]
    three_d = len(coords) == 3
    z = coords[2] if three_d else None
    ncx, ncy = len(x) - 1, len(y) - 1
    ncz = len(z) - 1 if three_d else 1
    r = {"faces": 0, "surface_measure": 0.0, "flux": 0.0, "traction": [0.0, 0.0, 0.0],
         "traction_magnitude": 0.0, "current": 0.0, "current_cells": 0}
    layers = [ncz // 2] if three_d else [0]
    for k in layers:
        for j in range(ncy):
            for i in range(ncx):
                c = i + ncx * (j + ncy * k)
                if positive[c]:
                    r["current"] += (x[i + 1] - x[i]) * (y[j + 1] - y[j]) * jz[c]
                    r["current_cells"] += 1
    samples = []
    for k in range(ncz):
        for j in range(ncy):
            for i in range(ncx):
                c = i + ncx * (j + ncy * k)
                if not in_region[c]:
                    continue
                if j + 1 < ncy and in_region[c + ncx]:
                    continue
                measure = (x[i + 1] - x[i]) * ((z[k + 1] - z[k]) if three_d else length_z)
                b, h = cell_b[c], cell_h[c]
                r["flux"] += measure * b[1]
                t = traction_on_positive_y_face(b, h)
                r["surface_measure"] += measure
                for a in range(3):
                    r["traction"][a] += measure * t[a]
                r["traction_magnitude"] += measure * math.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2])
                cz = 0.5 * (z[k] + z[k + 1]) if three_d else z_value
                samples.append((0.5 * (x[i] + x[i + 1]), cz, measure, t))
                r["faces"] += 1
    mirror_projection(samples, r)
    return r
