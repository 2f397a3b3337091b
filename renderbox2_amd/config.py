"""Readers for the reference's configuration files, mapped onto this
library's render kinds and parameters.

* Integrator files (``integrators/integrator_*.xml``), read as
  Configuration::get_*_integrator_params does
  (src/config/configuration.cpp:186-296): ``read_integrator_xml(path)``
  returns ``(kind, params)`` for ``SamplerRenderer(tracer, camera, kind,
  **params)`` / ``RenderGraph`` / ``RenderServer.submit_render``.
* The application file (``config.xml``, configuration.cpp:36-343):
  ``read_config_xml(path)`` returns the output size and file, the scene and
  BVH-cache files, the sampler-renderer mode / spp / iterations, the film
  filter, and the integrator (its file resolved beside ``config.xml`` or in
  an ``integrators/`` folder there, as the reference resolves it under
  ``src/integrators``).

Parameters the reference reads but never uses on the path are accepted and
ignored: the path tracer's ``usemaxdepth``, BDPT's ``mode`` (sk / mk / sort
are one wavefront pass here), ``preparation/numpaths`` (the cache has fixed
slots, no size estimate), ``lighttrace/rrstart`` (the light pass applies
roulette from the first bounce, integrator_bdpt_lvc_sk.cu:151-160) and
``sample_M``.
"""
import os
import xml.etree.ElementTree as ET

# Filter names of get_filter_params (configuration.cpp:298-343) that the
# film implements; "mitchell" and "sinc" parse there but have no filter.
_FILTERS = ("box", "triangle", "gaussian")


def _text(node, path, cast=str):
    e = node.find(path)
    if e is None or e.text is None:
        raise ValueError(f"missing <{path}>")
    return cast(e.text.strip())


def _raycast_shade(name):
    # configuration.cpp:228-248: three names, everything else is uniform
    return name if name in ("primitive_id", "material_id", "normals") else "uniform"


def read_integrator_xml(path):
    """(kind, params) of an integrator file.  kind is "ao", "pt", "bdpt" or
    "raycast" (the reference's type names ao / pt / bdpt_lvc / raycast,
    configuration.cpp:138-160)."""
    root = ET.parse(path).getroot()
    if root.tag != "integrator":
        raise ValueError(f"{path}: root element is <{root.tag}>, not <integrator>")
    kind = _text(root, "type")
    p = root.find("params")
    if p is None:
        raise ValueError(f"{path}: missing <params>")
    if kind == "pt":
        return "pt", dict(max_depth=_text(p, "maxdepth", int), rr_start=_text(p, "rrstart", int),
                          rr_prob=_text(p, "rrprob", float))
    if kind == "ao":
        return "ao", dict(ao_samples=_text(p, "samples", int), ao_radius=_text(p, "radius", float))
    if kind == "raycast":
        return "raycast", dict(shade=_raycast_shade(_text(p, "shade")))
    if kind == "bdpt_lvc":
        return "bdpt", dict(max_depth=_text(p, "cameratrace/maxdepth", int),
                            rr_start=_text(p, "cameratrace/rrstart", int),
                            rr_prob=_text(p, "cameratrace/rrprob", float),
                            connections=_text(p, "cameratrace/sample_N", int),
                            light_max_depth=_text(p, "lighttrace/maxdepth", int),
                            light_paths=_text(p, "lighttrace/numpaths", int))
    raise ValueError(f"{path}: unsupported integrator type {kind!r}")


def _filter_params(fp):
    name = _text(fp, "type")
    if name not in _FILTERS:
        raise ValueError(f"unsupported film filter {name!r} (box, triangle or gaussian)")
    out = dict(filter=name, filter_width=(_text(fp, "xwidth", float), _text(fp, "ywidth", float)))
    if name == "gaussian":
        out["filter_alpha"] = _text(fp, "alpha", float)
    return out


def _find_integrator_file(base, name):
    for cand in (os.path.join(base, name), os.path.join(base, "integrators", name)):
        if os.path.exists(cand):
            return cand
    raise FileNotFoundError(f"integrator file {name!r} not found beside the config or in integrators/")


def read_config_xml(path):
    """The application file as a dict:
    ``width``, ``height``, ``output`` (directory + file), ``scene_dir``,
    ``scene_file``, ``bvh_file`` and ``build_bvh`` (build, or load the
    cached tree), ``renderer`` (``mode`` "final" / "progressive", ``spp``,
    ``iterations``), ``kind`` and ``params`` (the integrator's, with the
    renderer's spp and the film filter merged in)."""
    root = ET.parse(path).getroot()
    if root.tag != "config":
        raise ValueError(f"{path}: root element is <{root.tag}>, not <config>")
    rp = root.find("renderer/params")
    if rp is None:
        raise ValueError(f"{path}: missing <renderer><params>")
    rtype = _text(root, "renderer/type")
    if rtype != "samplerrenderer":
        raise ValueError(f"{path}: unsupported renderer {rtype!r}")
    mode = _text(rp, "mode")
    # configuration.cpp:171-178 spells the second mode "progresive"
    renderer = dict(mode="progressive" if mode in ("progressive", "progresive") else "final",
                    spp=_text(rp, "spp", int), iterations=_text(rp, "iterations", int))
    fp = rp.find("filter/params")
    film = _filter_params(fp) if fp is not None else {}
    ifile = _find_integrator_file(os.path.dirname(os.path.abspath(path)), _text(root, "integrator/file"))
    kind, params = read_integrator_xml(ifile)
    params = dict(params, spp=renderer["spp"], **film)
    scene_dir = _text(root, "directories/scenefilepath")
    out_dir = root.find("directories/output")
    out_dir = out_dir.text.strip() if out_dir is not None and out_dir.text else ""
    return dict(width=_text(root, "output/width", int), height=_text(root, "output/height", int),
                output=out_dir + _text(root, "output/file"),
                scene_dir=scene_dir, scene_file=_text(root, "scene/file"),
                bvh_file=scene_dir + _text(root, "scene/cudabvh/file"),
                build_bvh=_text(root, "scene/cudabvh/build").lower() == "yes",
                renderer=renderer, kind=kind, params=params, integrator_file=ifile)
