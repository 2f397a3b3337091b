"""The reference's configuration files (src/config/configuration.cpp:36-343):
the application file and the per-integrator files, read into render kinds
and parameters.  The fixtures are written here in the files' own schema
(the element names the reference's reader looks up); expectations come from
the file text."""
import pytest

from renderbox2_amd import _lib as L
from renderbox2_amd.config import read_config_xml, read_integrator_xml

PT_XML = """<!-- Integrator File : Path Tracing integrator -->
<integrator>
    <type>pt</type>
    <params>
        <maxdepth>10</maxdepth>
        <usemaxdepth>yes</usemaxdepth>
        <rrstart>5</rrstart>
        <rrprob>0.5</rrprob>
    </params>
</integrator>"""

AO_XML = """<integrator><type>ao</type>
    <params><samples>32</samples><radius>1.0</radius></params></integrator>"""

BDPT_XML = """<integrator>
    <type>bdpt_lvc</type>
    <params>
        <mode>mk</mode>
        <preparation><numpaths>10000</numpaths></preparation>
        <lighttrace><numpaths>100000</numpaths><maxdepth>8</maxdepth><rrstart>4</rrstart></lighttrace>
        <cameratrace>
            <maxdepth>8</maxdepth><rrstart>4</rrstart><rrprob>0.3</rrprob>
            <sample_M>10</sample_M><sample_N>5</sample_N>
        </cameratrace>
    </params>
</integrator>"""


def raycast_xml(shade):
    return f"<integrator><type>raycast</type><params><shade>{shade}</shade></params></integrator>"


CONFIG_XML = """<!-- Config File Version 2.0 -->
<config>
    <directories>
        <output>out/</output>
        <scenefilepath>scenes/cornell/</scenefilepath>
    </directories>
    <flags><render>yes</render><gui>no</gui></flags>
    <output><width>800</width><height>600</height><file>test_dl.jpg</file></output>
    <scene>
        <file>cornell_box_cubes.xml</file>
        <cudabvh><file>cornell_box_cubes.bvh</file><build>no</build></cudabvh>
    </scene>
    <renderer>
        <type>samplerrenderer</type>
        <params>
            <mode>{mode}</mode>
            <spp>50</spp>
            <iterations>3</iterations>
            <filter><params>{filter}</params></filter>
        </params>
    </renderer>
    <integrator><file>{integrator}</file></integrator>
</config>"""


def write(tmp_path, name, text):
    p = tmp_path / name
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(text)
    return p


def test_integrator_files(tmp_path):
    assert read_integrator_xml(write(tmp_path, "pt.xml", PT_XML)) == \
        ("pt", dict(max_depth=10, rr_start=5, rr_prob=0.5))
    assert read_integrator_xml(write(tmp_path, "ao.xml", AO_XML)) == ("ao", dict(ao_samples=32, ao_radius=1.0))
    kind, p = read_integrator_xml(write(tmp_path, "bd.xml", BDPT_XML))
    assert kind == "bdpt"
    assert p == dict(max_depth=8, rr_start=4, rr_prob=0.3, connections=5, light_max_depth=8, light_paths=100000)


@pytest.mark.parametrize("name,want", [("primitive_id", "primitive_id"), ("material_id", "material_id"),
                                       ("normals", "normals"), ("uvs", "uniform"), ("anything", "uniform")])
def test_raycast_shade_names_follow_the_reference_reader(tmp_path, name, want):
    # configuration.cpp:228-248 knows three names; every other one is uniform
    assert read_integrator_xml(write(tmp_path, "rc.xml", raycast_xml(name))) == ("raycast", dict(shade=want))


def test_integrator_file_errors(tmp_path):
    with pytest.raises(ValueError, match="unsupported integrator"):
        read_integrator_xml(write(tmp_path, "x.xml", "<integrator><type>photon</type><params/></integrator>"))
    with pytest.raises(ValueError, match="maxdepth"):
        read_integrator_xml(write(tmp_path, "y.xml", "<integrator><type>pt</type><params/></integrator>"))
    with pytest.raises(ValueError, match="not <integrator>"):
        read_integrator_xml(write(tmp_path, "z.xml", "<config/>"))


def test_config_file(tmp_path):
    write(tmp_path, "integrators/integrator_pt.xml", PT_XML)
    cfg = read_config_xml(write(tmp_path, "config.xml", CONFIG_XML.format(
        mode="final", integrator="integrator_pt.xml",
        filter="<type>gaussian</type><xwidth>2.0</xwidth><ywidth>1.5</ywidth><alpha>2.5</alpha>")))
    assert (cfg["width"], cfg["height"]) == (800, 600)
    assert cfg["output"] == "out/test_dl.jpg"
    assert cfg["scene_file"] == "cornell_box_cubes.xml" and cfg["scene_dir"] == "scenes/cornell/"
    assert cfg["bvh_file"] == "scenes/cornell/cornell_box_cubes.bvh" and cfg["build_bvh"] is False
    assert cfg["renderer"] == dict(mode="final", spp=50, iterations=3)
    assert cfg["kind"] == "pt"
    assert cfg["params"] == dict(max_depth=10, rr_start=5, rr_prob=0.5, spp=50, filter="gaussian",
                                 filter_width=(2.0, 1.5), filter_alpha=2.5)


def test_config_params_reach_render_params(tmp_path, rbm):
    write(tmp_path, "integrator_bdpt_lvc.xml", BDPT_XML)
    cfg = read_config_xml(write(tmp_path, "config.xml", CONFIG_XML.format(
        mode="progresive", integrator="integrator_bdpt_lvc.xml",
        filter="<type>triangle</type><xwidth>1.0</xwidth><ywidth>1.0</ywidth>")))
    assert cfg["renderer"]["mode"] == "progressive"
    p = rbm.render_params(**cfg["params"])
    assert (p.spp, p.max_depth, p.rr_start, p.connections, p.light_max_depth, p.light_paths) == (50, 8, 4, 5, 8, 100000)
    assert abs(p.rr_prob - 0.3) < 1e-7
    assert p.filter == L.RB_FILTER_TRIANGLE and tuple(p.filter_width) == (1.0, 1.0)


def test_config_rejects_filters_the_film_lacks(tmp_path):
    write(tmp_path, "integrator_pt.xml", PT_XML)
    with pytest.raises(ValueError, match="mitchell"):
        read_config_xml(write(tmp_path, "config.xml", CONFIG_XML.format(
            mode="final", integrator="integrator_pt.xml",
            filter="<type>mitchell</type><xwidth>2</xwidth><ywidth>2</ywidth>")))
    with pytest.raises(FileNotFoundError):
        read_config_xml(write(tmp_path, "c2.xml", CONFIG_XML.format(
            mode="final", integrator="missing.xml", filter="<type>box</type><xwidth>0.5</xwidth><ywidth>0.5</ywidth>")))
