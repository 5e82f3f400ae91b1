"""Input adapters (host code, no GPU): the calibration loaders of
kinect.cpp:78-161 through the library's parser, checked against Python's
own XML / JSON readers on committed fixtures and, where the reference tree
is mounted, on its resources/infrared.xml / .json; recorded-depth files."""
import json
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oglfusionkinectv2_amd import io

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = "/root/reference/resources"


def xml_expect(path):
    root = ET.fromstring("<r>" + open(path).read().replace('<?xml version="1.0"?>', "") + "</r>")
    cm = [float(v) for v in root.find(".//Camera_Matrix/data").text.replace(",", " ").split()]
    di = [float(v) for v in root.find(".//Distortion_Coefficients/data").text.replace(",", " ").split()]
    return cm, di


def check(p, cm, di):
    assert p.fx == np.float32(cm[0]) and p.fy == np.float32(cm[4])
    assert p.ppx == np.float32(cm[2]) and p.ppy == np.float32(cm[5])
    np.testing.assert_array_equal([p.k1, p.k2, p.p1, p.p2, p.k3], np.float32(di[:5]))


def test_xml_calibration():
    rc, p = io.loadCalibration(os.path.join(GOLD, "calib_ir.xml"))
    assert rc == 1
    cm, di = xml_expect(os.path.join(GOLD, "calib_ir.xml"))
    check(p, cm, di)
    assert (p.fx, p.fy, p.ppx, p.ppy) == (365.125, 365.9375, 256.1875, 210.4375)
    assert p.depth_cam() == (365.125, 365.125, 256.1875, 210.4375)


def test_json_calibration():
    rc, p = io.loadCalibration(os.path.join(GOLD, "calib_ir.json"))
    assert rc == 1
    j = json.load(open(os.path.join(GOLD, "calib_ir.json")))
    cm = [v for row in j["camera_matrix"] for v in row]
    check(p, cm, j["dist_coeff"])


def test_other_extensions_and_bad_files(tmp_path):
    assert io.loadCalibration(str(tmp_path / "calib.yml")) == (0, None)
    bad = tmp_path / "bad.xml"
    bad.write_text("<Camera_Matrix><rows>3</rows></Camera_Matrix>")
    from oglfusionkinectv2_amd._lib import KfError
    with pytest.raises(KfError):
        io.loadCalibration(str(bad))
    with pytest.raises(KfError):
        io.loadCalibration(str(tmp_path / "missing.json"))


@pytest.mark.skipif(not os.path.exists(REF), reason="reference tree not mounted")
def test_reference_resources_parse():
    rc, p = io.loadCalibration(os.path.join(REF, "infrared.xml"))
    assert rc == 1
    check(p, *xml_expect(os.path.join(REF, "infrared.xml")))
    rc, q = io.loadCalibration(os.path.join(REF, "infrared.json"))
    assert rc == 1
    j = json.load(open(os.path.join(REF, "infrared.json")))
    check(q, [v for row in j["camera_matrix"] for v in row], j["dist_coeff"])


def test_recorded_depth_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    for dt in (np.float32, np.uint16):
        frames = (rng.uniform(0, 4500, (3, 424, 512))).astype(dt)
        path = str(tmp_path / ("d_%s.raw" % np.dtype(dt).name))
        io.write_depth_frames(path, frames)
        back = io.read_depth_frames(path, dtype=dt)
        assert back.shape == (3, 424, 512)
        np.testing.assert_array_equal(back, frames)
