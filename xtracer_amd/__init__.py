"""xtracer_amd: MI355X-native ray-traversal and integrator path of xtracer's xtcore.

Host API (scene loading, octree/kd-tree builds) and device API (closest-hit
queries, renders, serving) over the C ABI in include/xtracer_amd.h.
"""
from ._native import (ACCEL_KDTREE, ACCEL_OCTREE, AA_GRID, AA_RANDOM, ACCUM_DEFAULT, ACCUM_FP64,  # noqa: F401
                      ACCUM_REFERENCE, AO, DEPTH, EMISSION, NORMAL, UV, WIREFRAME, HIT_NONE, HIT_PLANE_BASE, INTEGRATORS, PATHTRACER, STENCIL,
                      TUNING_AUTO, Tuning, SERVE_FAIR, SERVE_SRPT,
                      Device, Job, JobStats, Scene, XtError, lib, make_camera, make_job, write_image, read_image)

__all__ = ["Scene", "Device", "Job", "JobStats", "XtError", "make_job", "make_camera", "write_image", "read_image", "lib"]
