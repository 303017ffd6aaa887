#!/usr/bin/env python3
"""Extract the hop table of the reference's generated displacement routine.

Reads get_diff_sites_displacement from the proclist_acf.f90 files the
reference ships as test fixtures (tests/test_acf/_tmp_export_<backend>/
ref_src, generated for tests/test_acf/2d_grid.xml, copied here as
tests/golden/2d_grid.xml) and writes, as JSON data, per process the site
the particle leaves (lsite_old) and the site it reaches (lsite_new), each
as [dx, dy, dz, site]: for local_smart relative to the chosen site, `site`
being the site name or "exec" for the chosen site's own type
(ref_hops_2d_grid.json); for lat_int relative to the chosen site's cell,
whose site 1 holds the process (ref_hops_2d_grid_lat_int.json).

Usage: python tests/golden/extract_reference_hops.py [--ref /root/reference]
"""
import argparse
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "tests/test_acf/_tmp_export_%s/ref_src/proclist_acf.f90"
OUT = {"local_smart": "ref_hops_2d_grid.json", "lat_int": "ref_hops_2d_grid_lat_int.json"}
LAYER = "default"


def parse(text):
    body = text.split("subroutine get_diff_sites_displacement", 1)[1]
    body = body.split("end subroutine get_diff_sites_displacement", 1)[0]
    out = {}
    proc = None
    for line in body.splitlines():
        line = line.strip()
        m = re.match(r"case\((\w+)\)", line)
        if m:
            proc = m.group(1)
            out[proc] = {}
            continue
        # local_smart writes "(lsite + (/.../))", lat_int "lsite + (/.../)"
        m = re.match(r"lsite_(new|old) = \(?lsite(?: \+ \(/(.*)/\))?\)?$", line)
        if m and proc:
            which = "fin" if m.group(1) == "new" else "init"
            if m.group(2) is None:
                out[proc][which] = [0, 0, 0, "exec"]
            else:
                parts = [p.strip() for p in m.group(2).split(",")]
                dx, dy, dz = (int(v) for v in parts[:3])
                n = parts[3]
                if n == "0":
                    site = "exec"
                else:
                    first = n.split("-")[0].strip()
                    assert first.startswith(LAYER + "_"), n
                    site = first[len(LAYER) + 1:]
                out[proc][which] = [dx, dy, dz, site]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    for backend, name in OUT.items():
        with open(os.path.join(a.ref, SRC % backend)) as f:
            table = parse(f.read())
        out = os.path.join(HERE, name)
        with open(out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        print(out, len(table))


if __name__ == "__main__":
    main()
