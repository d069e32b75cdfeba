"""MI355X (gfx950) kD-tree Barnes-Hut step behind the reference's
simple_sim / build_tree / calc_accel API.  See include/kdtree_hip.h for the C
ABI; this package is the Python host interface over it."""
from .kdtree import (KD_LAYOUT_COMPACT, KD_LAYOUT_PADDED, KD_PREC_FP64,
                     KD_PREC_MIXED, KD_PREC_MIXED_D64, Simulation, Tree,
                     allocate_node_vec, build_tree, calc_accel,
                     circular_orbits, ring, simple_sim, two_bodies)

__all__ = ["Simulation", "Tree", "allocate_node_vec", "build_tree", "calc_accel",
           "circular_orbits", "ring", "simple_sim", "two_bodies",
           "KD_PREC_FP64", "KD_PREC_MIXED", "KD_PREC_MIXED_D64",
           "KD_LAYOUT_COMPACT", "KD_LAYOUT_PADDED"]
