"""Test-data generators of the reference's test suite, restated in numpy.

Each function follows helpers/GeometryGeneration.cpp of GPU-Voxels (paths
relative to packages/gpu_voxels/src/gpu_voxels/), including its float32
accumulation, so the known answers of the reference's tests apply unchanged.
"""
import numpy as np

# ArgsFixture defaults, test/testing_fixtures.hpp:137-179
DIM_X, DIM_Y, DIM_Z = 89, 123, 74
NUM_POINTS = 10000


def create_box_of_points(bmin, bmax, delta):
    """createBoxOfPoints(min, max, delta), GeometryGeneration.cpp:90-106 (float loop counters)."""
    f = np.float32
    out = []
    x = f(bmin[0])
    while x <= f(bmax[0]):
        y = f(bmin[1])
        while y <= f(bmax[1]):
            z = f(bmin[2])
            while z <= f(bmax[2]):
                out.append((x, y, z))
                z = f(z + f(delta))
            y = f(y + f(delta))
        x = f(x + f(delta))
    return np.array(out, dtype=np.float32).reshape(-1, 3)


def create_equidistant_points_in_box(max_points, dims, side):
    """createEquidistantPointsInBox, GeometryGeneration.cpp:237-265."""
    pts = []
    f = np.float32
    s = f(side)
    for i in range((dims[0] - 1) // 2):
        for j in range((dims[1] - 1) // 2):
            for k in range((dims[2] - 1) // 2):
                if len(pts) >= max_points:
                    return np.array(pts, np.float32)
                # x = i * 2 * side + side / 2.0  (float * int, then + double -> float)
                pts.append((f(np.float64(f(i * 2 * s)) + np.float64(s) / 2.0),
                            f(np.float64(f(j * 2 * s)) + np.float64(s) / 2.0),
                            f(np.float64(f(k * 2 * s)) + np.float64(s) / 2.0)))
    return np.array(pts, np.float32)


def create_checkerboard(max_points, dims, side):
    """createNonOverlapping3dCheckerboard, GeometryGeneration.cpp:268-299."""
    black, white = [], []
    f = np.float32
    s = f(side)
    for i in range((dims[0] - 1) // 2):
        for j in range((dims[1] - 1) // 2):
            for k in range((dims[2] - 1) // 2):
                if len(black) >= max_points:
                    return np.array(black, np.float32), np.array(white, np.float32)
                black.append((f(np.float64(f(i * 2 * s)) + s / 2.0), f(np.float64(f(j * 2 * s)) + s / 2.0),
                              f(np.float64(f(k * 2 * s)) + s / 2.0)))
                white.append((f(np.float64(f((i * 2 + 1) * s)) + s / 2.0),
                              f(np.float64(f((j * 2 + 1) * s)) + s / 2.0),
                              f(np.float64(f((k * 2 + 1) * s)) + s / 2.0)))
    return np.array(black, np.float32), np.array(white, np.float32)


def bitvector_stream(bits32):
    """operator<< of BitVector (helpers/BitVector.h:271-288): bit 0 first."""
    return "".join(format(int(b), "08b")[::-1] for b in bits32)


def set_bits(*idx):
    b = np.zeros(32, np.uint8)
    for i in idx:
        b[i >> 3] |= 1 << (i & 7)
    return b


def bits_of(mask):
    m = np.asarray(mask, np.uint8)
    return [i for i in range(256) if (m[i >> 3] >> (i & 7)) & 1]
