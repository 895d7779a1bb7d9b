"""Text dumps of the reference driver (src/helperFunctions.cu:76-169), same
file formats: results/pointCloud.txt, results/kdtree.txt,
results/hmatrixstructure.txt.  Inputs are host arrays (numpy); see
hmx.KDTree.copy and hmx.Structure.tiles for getting them off the device.
"""
import numpy as np


def _compact(x):
    x = x & 0x55555555
    x = (x | (x >> 1)) & 0x33333333
    x = (x | (x >> 2)) & 0x0F0F0F0F
    x = (x | (x >> 4)) & 0x00FF00FF
    x = (x | (x >> 8)) & 0x0000FFFF
    return x


def point_cloud(path, points, n, d):
    """printPointCloud (src/helperFunctions.cu:94-115): one line per dimension,
    labelled "x" for the first and "y" for every other, SoA input p[j*n+i]."""
    p = np.asarray(points, dtype=np.float64).reshape(d, n)
    with open(path, "w") as f:
        for i in range(d):
            f.write("x = [" if i == 0 else "y = [")
            f.write("".join("%f, " % v for v in p[i]))
            f.write("]\n")


def kdtree(path, n, d, bucket_size, perm, offsets, points):
    """printKDTree (src/helperFunctions.cu:117-169), FULL_TREE division:
    permutation, segment offsets, then the coordinates bucket by bucket for
    the first n/bucket_size buckets of bucket_size points."""
    p = np.asarray(points, dtype=np.float64).reshape(d, n)
    perm = np.asarray(perm)
    with open(path, "w") as f:
        f.write("segmentIndices\n")
        f.write("".join("%d " % v for v in perm[:n]))
        f.write("\n\n")
        f.write("segmentOffsets\n")
        f.write("".join("%d " % v for v in offsets))
        f.write("\n\n\n")
        f.write("n = %d\n" % n)
        f.write("bs = %d\n" % bucket_size)
        for i in range(d):
            f.write("X = [" if i == 0 else "Y = [")
            for j in range(n // bucket_size):
                f.write("[ ")
                f.write("".join("%f, " % p[i, perm[j * bucket_size + k]] for k in range(bucket_size)))
                f.write("], ")
            f.write("]\n")


def structure(path, num_levels, tiles_per_level):
    """printMatrixStructure (src/helperFunctions.cu:76-92): per level l the
    tiles as column-major indices col * 2^l + row."""
    with open(path, "w") as f:
        f.write("%d\n" % num_levels)
        f.write("[ ")
        for i in range(num_levels - 1):
            nt = 1 << (i + 1)
            f.write("[ ")
            for code in np.asarray(tiles_per_level[i], dtype=np.int64):
                x, y = _compact(int(code)), _compact(int(code) >> 1)
                f.write("%d, " % (x * nt + y))
            f.write(" ], ")
        f.write(" ]")
