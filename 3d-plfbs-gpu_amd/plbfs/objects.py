"""Lagrangian objects for the immersed boundary.

rigid_cylinder follows the reference's node layout (src/lagrangian_object.cpp:79-113,
node area 2 pi r / n * depth as in src/immersed_boundary_method.cu:168).
rigid_sphere places n nodes on a Fibonacci (golden-angle) lattice; node area
4 pi r^2 / n.  csrc/host/case.cpp builds the same node sets from XML.
"""
import math

import numpy as np


def fibonacci_sphere(n, centre, radius):
    ga = math.pi * (3.0 - math.sqrt(5.0))
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    ph = ga * i
    x = np.stack([centre[0] + radius * r * np.cos(ph), centre[1] + radius * r * np.sin(ph), centre[2] + radius * z], 1)
    return x, 4.0 * math.pi * radius * radius / n


def rigid_cylinder(num_nodes, centre, radius, depth):
    dn = int(math.sqrt(num_nodes))
    dz = depth / dn
    pts = []
    for j in range(dn):
        for i in range(dn):
            pts.append((centre[0] + radius * math.sin(2 * math.pi * i / dn),
                        centre[1] + radius * math.cos(2 * math.pi * i / dn),
                        centre[2] + depth / 2 - j * depth / dn - dz / 2))
    return np.array(pts), 2 * math.pi * radius / num_nodes * depth


def sphere_nodes_for_diameter(d_cells):
    """About one node per h^2 of sphere surface: round(pi * D^2) for D in cells."""
    return int(round(math.pi * d_cells * d_cells))
