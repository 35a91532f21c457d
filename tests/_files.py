"""Writers for the reference's stage-boundary file set (build/output/proc/)."""
import struct

import numpy as np


def write_bin2d_cols(path, root):
    """rootData_complex.bin: dims [S, N], payload flat r + p*S (root is [N][S])."""
    N, S = root.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<III", 2, 1, 6) + b"double".ljust(16, b"\0") + struct.pack("<QQ", S, N))
        f.write(np.ascontiguousarray(root, dtype=np.complex128).tobytes())


def write_vec(path, v):
    with open(path, "w") as f:
        f.write(",".join("%.6f" % x for x in v) + "\n")


def write_mat(path, m):
    with open(path, "w") as f:
        for row in m:
            f.write(",".join("%.6f" % x for x in row) + "\n")


def write_proc_dir(d, g):
    """Input set of a ref_bp_*.npz golden case into directory d."""
    d.mkdir(parents=True, exist_ok=True)
    write_bin2d_cols(d / "rootData_complex.bin", g["root"])
    write_vec(d / "grid_xAxis.csv", g["gx"])
    write_vec(d / "grid_yAxis.csv", g["gy"])
    write_vec(d / "grid_zAxis.csv", g["gz"])
    write_vec(d / "DEM_X_vec.csv", g["dem_x"])
    write_vec(d / "DEM_Y_vec.csv", g["dem_y"])
    write_mat(d / "DEM_Z_matrix.csv", g["dem_z"])
    write_mat(d / "Tx_positions.csv", g["tx"])
    write_mat(d / "Rx_positions.csv", g["rx"])
    (d / "environment.txt").write_text("n1=%r\nn2=%r\n" % (float(g["n1"]), float(g["n2"])))
