"""Write the golden fixtures of the tests from data files the reference holds.

Run in a container where the reference is mounted at /root/reference:
    python tests/golden/make_fixtures.py
Reads only data files (no reference code is imported or executed):
  pynicamdc/test/case2/prepdata/rl01-prc000008.toml     region link table, rlevel 1
  pynicamdc/test/case2/json_gl05rl01pe08/bboundary_GL05RL01.pe0000000[0-7].json
                                                         gl05 grid points incl. halos
  pynicamdc/test/case2/json_gl05rl01pe08/vgrid40_stretch_45km.json
                                                         vertical grid (full and half levels)
Writes:
  rl01_region_links.json   {"SW"|"NW"|"NE"|"SE": [[rgn, dir], ...] for rgnid 0..39}
  gl05_grd_x.npz           grd_x (40, 18, 18, 3) indexed (l, j, i, xyz), unit sphere
  vgrid40_stretch_45km.json  copy of the two level arrays (gz, gzh)
"""
import json
import os
import re

import numpy as np

REF = "/root/reference/pynicamdc/test/case2"
OUT = os.path.dirname(os.path.abspath(__file__))


def region_links():
    text = open(os.path.join(REF, "prepdata", "rl01-prc000008.toml")).read()
    blocks = re.split(r"\[RGN_LINK_INFO\.\d+\]", text)[1:]
    links = {d: [] for d in ("SW", "NW", "NE", "SE")}
    for n, b in enumerate(blocks):
        rid = int(re.search(r"RGNID\s*=\s*(\d+)", b).group(1))
        assert rid == n
        for d in links:
            m = re.search(rf"{d}\s*=\s*\[\s*(\d+)\s*,\s*(\d+)\s*,?\s*\]", b)
            links[d].append([int(m.group(1)), int(m.group(2))])
    mng = re.findall(r"MNG_RGNID\s*=\s*\[([^\]]*)\]", text)
    order = [int(x) for blk in mng for x in blk.replace(",", " ").split()]
    assert order == list(range(40))
    with open(os.path.join(OUT, "rl01_region_links.json"), "w") as f:
        json.dump(links, f)


def gl05_grid():
    g = 18
    xs = []
    for pe in range(8):
        d = json.load(open(os.path.join(REF, "json_gl05rl01pe08", f"bboundary_GL05RL01.pe{pe:08d}.json")))
        a = np.stack([np.array(d[f"grd_x_{c}"], dtype=np.float64) for c in "xyz"], axis=-1)   # (5, g*g, 3)
        xs.append(a.reshape(5, g, g, 3))     # flat index j*g + i -> (l, j, i)
    np.savez_compressed(os.path.join(OUT, "gl05_grd_x.npz"), grd_x=np.concatenate(xs, axis=0))


def vgrid():
    d = json.load(open(os.path.join(REF, "json_gl05rl01pe08", "vgrid40_stretch_45km.json")))
    with open(os.path.join(OUT, "vgrid40_stretch_45km.json"), "w") as f:
        json.dump({"gz": d["set1"], "gzh": d["set2"]}, f)


if __name__ == "__main__":
    region_links()
    gl05_grid()
    vgrid()
