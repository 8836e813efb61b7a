"""Extract the Panda sphere model and base transforms from the reference's
robot headers (read as text) into ``panda_tables.json``.

The reference keeps its robot model as numeric ``__constant__`` tables in
src/robots/panda.cuh:16-595 (single arm) and
src/robots/panda_multi_geometry.cuh:17-516 (the same values under ``panda2_``
names for the multi-arm checkers); the 4- and 5-arm base transforms live in
src/robots/panda.cuh:16-73 and src/robots/panda_{four,five}.cuh:17-47.
This script parses those tables only (numbers, not code) and writes them as a
JSON fixture that the CPU tests compare against the build's own geometry
header.  Run it in a checkout where the reference is mounted:

    python tests/golden/gen_panda_tables.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?"


def table(text, name):
    m = re.search(r"__constant__\s+\w+\s+" + re.escape(name) + r"\s*(?:\[[^\]]*\])*\s*=\s*\{(.*?)\};",
                  text, re.S)
    if not m:
        raise KeyError(name)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(x.rstrip("f")) for x in re.findall(NUM, body)]


def main(ref):
    ref = Path(ref)
    single = (ref / "src/robots/panda.cuh").read_text()
    multi = (ref / "src/robots/panda_multi_geometry.cuh").read_text()
    four = (ref / "src/robots/panda_four.cuh").read_text()
    five = (ref / "src/robots/panda_five.cuh").read_text()
    robots = (ref / "src/planning/Robots.hh").read_text()

    out = {}
    for key, suffix in [("coarse_spheres", "approx_spheres_array"),
                        ("coarse_fixed_transforms", "approx_fixed_transforms"),
                        ("coarse_sphere_to_joint", "approx_sphere_to_joint"),
                        ("coarse_joint_to_sphere_count", "approx_joint_to_sphere_count"),
                        ("coarse_joint_types", "approx_joint_types"),
                        ("coarse_self_cc_ranges", "approx_self_cc_ranges"),
                        ("coarse_joint_parents", "approx_joint_parents"),
                        ("coarse_joint_id_to_dof", "approx_joint_id_to_dof"),
                        ("fine_spheres", "spheres_array"),
                        ("fine_fixed_transforms", "fixed_transforms"),
                        ("fine_sphere_to_joint", "sphere_to_joint"),
                        ("fine_joint_to_sphere_count", "joint_to_sphere_count"),
                        ("fine_joint_types", "joint_types"),
                        ("fine_self_cc_ranges", "self_cc_ranges"),
                        ("fine_joint_parents", "joint_parents"),
                        ("fine_joint_id_to_dof", "joint_id_to_dof")]:
        a = table(single, "panda_" + suffix)
        b = table(multi, "panda2_" + suffix)
        if a != b:
            raise SystemExit(f"single-arm and multi-arm tables differ: {suffix}")
        out[key] = a
    out["base_4"] = table(single, "quad_franka_base_transform")
    out["base_5"] = table(single, "five_franka_base_transform")
    if table(four, "robot_base_transform_4panda") != out["base_4"]:
        raise SystemExit("4-arm base tables differ")
    if table(five, "robot_base_transform_5panda") != out["base_5"]:
        raise SystemExit("5-arm base tables differ")
    m = re.search(r"struct Panda \{.*?get_s_m.*?\{(.*?)\};.*?get_s_a.*?\{(.*?)\};", robots, re.S)
    out["scale_s_m"] = [float(x.rstrip("f")) for x in re.findall(NUM, m.group(1))]
    out["scale_s_a"] = [float(x.rstrip("f")) for x in re.findall(NUM, m.group(2))]
    dst = Path(__file__).with_name("panda_tables.json")
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", dst, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
