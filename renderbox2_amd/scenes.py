"""Seeded synthetic scenes and cameras of the benchmark configurations.

Exact counts (the configuration texts are internally inconsistent; these are
the explicit choices, recorded with every result):
  ao_room      room (2,048 tris, [-5,5]^3) + 60 icospheres s3 (76,800)  = 78,848 tris
  trace_1m     750 icospheres s3 (960,000) + 20,000 thin quads (40,000)  = 1,000,000 tris
  pt_room      room (2,048) + 187 icospheres s4 + 813 s3 (1,998,080)
               + 4 emissive quads (8)                                     = 2,000,136 tris
  trace_scaled(T): round(0.96 T / 1280) icospheres s3 + the rest as thin quads = T tris
  pt_scaled(T) : pt_room with the sphere count (and s4 share) set so the
                 total is within one sphere of T; pt_scaled(2e6) = pt_room
"""
from . import _lib as L
from .core import BVH, Platform, PerspectiveCamera, Scene

ICO3 = 1280
ICO4 = 5120


def ao_room(seed=0, smooth=False):
    """smooth: the icospheres carry their analytic vertex normals (same geometry)."""
    s = Scene()
    m = s.add_diffuse_bsdf((0.7, 0.7, 0.7))
    s.generate(L.RB_GEN_ROOM, c=(-5.0, 5.0), seed=seed, material_id=m)
    s.generate(L.RB_GEN_ICOSPHERES, count=60, subdiv=3, seed=seed, r=(0.1, 0.5), c=(-4.0, 4.0), material_id=m,
               smooth=smooth)
    return s


def ao_camera(width=512, height=512):
    return PerspectiveCamera((0.0, 0.5, -4.8), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (width, height), 70.0)


def trace_scaled(total_tris, seed=0):
    """Icospheres (s3, radius U[0.2,1.0]) + thin 1:50 quads (length U[0.5,2]) in [-10,10]^3."""
    spheres = int(round(0.96 * total_tris / ICO3))
    rest = total_tris - spheres * ICO3
    if rest < 0 or rest % 2:
        spheres -= 1
        rest = total_tris - spheres * ICO3
    quads = rest // 2
    s = Scene()
    m = s.add_diffuse_bsdf((0.5, 0.5, 0.5))
    s.generate(L.RB_GEN_ICOSPHERES, count=spheres, subdiv=3, seed=seed, r=(0.2, 1.0), c=(-9.0, 9.0), material_id=m)
    s.generate(L.RB_GEN_THIN_QUADS, count=quads, seed=seed, r=(0.5, 2.0), c=(-9.5, 9.5), material_id=m)
    return s


def trace_1m(seed=0):
    return trace_scaled(1_000_000, seed)


def trace_camera(width=2048, height=1024):
    return PerspectiveCamera((0.0, 4.0, -28.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (width, height), 50.0)


def pt_room(seed=0, spheres_hi=187, spheres_lo=813, emit=(12.0, 12.0, 12.0), smooth=False):
    s = Scene()
    wall = s.add_diffuse_bsdf((0.6, 0.6, 0.6))
    s.generate(L.RB_GEN_ROOM, c=(-5.0, 5.0), seed=seed, material_id=wall)
    s.generate(L.RB_GEN_ICOSPHERES, count=spheres_hi + spheres_lo, subdiv=3, subdiv_hi=4, hi_count=spheres_hi,
               seed=seed, r=(0.1, 0.5), c=(-4.0, 4.0), mix_materials=True, smooth=smooth)
    s.generate(L.RB_GEN_EMITTER_QUADS, count=4, seed=seed, r=(0.25, 0.0), c=(-5.0, 5.0), emit=emit)
    return s


def pt_scaled(total_tris, seed=0):
    """The path-tracing room with about total_tris triangles: up to 1000
    spheres of subdivision 3, then s4 spheres replacing s3 ones."""
    fixed = 2048 + 8
    lo = int(round((total_tris - fixed) / ICO3))
    if lo <= 1000:
        return pt_room(seed=seed, spheres_hi=0, spheres_lo=max(lo, 1))
    hi = int(round((total_tris - fixed - 1000 * ICO3) / (ICO4 - ICO3)))
    hi = min(hi, 1000)
    return pt_room(seed=seed, spheres_hi=hi, spheres_lo=1000 - hi)


def pt_camera(width=1024, height=1024):
    return PerspectiveCamera((0.0, 0.0, -4.8), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (width, height), 70.0)


def build(scene, max_leaf=None, threads=0):
    p = Platform()
    if max_leaf:
        p.set_leaf_preferences(1, max_leaf)
    from .core import BuildParams
    return BVH(scene, p, BuildParams(threads=threads))
