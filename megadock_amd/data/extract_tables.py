"""Extract MEGADOCK's data tables (rotation sets, CHARMM radii/charges, ACE values)
from the reference's ``src/parameter.cpp`` into plain text data files.

Tables copied as data (they must be identical for results to match):
  * rotation sets     src/parameter.cpp:3793-3810 (1), :3812-3830 (3), :3833-3873 (24),
                      :3876-4251 (360), :4254-7869 (3600), :7872-61886 (54000)
  * CHARMM radii      src/parameter.cpp:473-1391  (``_Charmmr.insert``)
  * CHARMM charges    src/parameter.cpp:1394-2312 (``_Charmmc.insert``)
  * ACE values        src/parameter.cpp:2315-3790 (``_ACE.insert``)

The literal text of every number is kept, so loaders can reproduce the reference's
``double literal -> float`` conversion exactly.  Key order and duplicates are kept
too; ``std::map::insert`` keeps the FIRST value for a duplicated key, and every
loader in this package applies the same rule.

Usage:  python extract_tables.py /root/reference/src/parameter.cpp
The outputs are committed; this script only documents their provenance.
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

ROT_FUNCS = {
    "dangle_rot1": 1, "dangle_rot3": 3, "dangle_rot24": 24,
    "dangle_rot360": 360, "dangle_rot3600": 3600, "dangle_rot54000": 54000,
}
MAP_TABLES = {"_Charmmr": "charmm_radius.tsv", "_Charmmc": "charmm_charge.tsv",
              "_ACE": "ace.tsv"}


def main(path):
    text = open(path).read().split("\n")
    # rotation sets
    cur = None
    rows = []
    for line in text:
        m = re.match(r"void Parameter::(dangle_rot\d+)\(\)", line)
        if m:
            cur = m.group(1)
            rows = []
            continue
        if cur is None:
            continue
        s = line.strip()
        if s.startswith("};"):
            n = ROT_FUNCS[cur]
            assert len(rows) == n, (cur, len(rows))
            with open(os.path.join(HERE, "rotations_%d.txt" % n), "w") as f:
                f.write("# ZXZ Euler angles (rad), phi theta psi; from src/parameter.cpp %s()\n" % cur)
                for r in rows:
                    f.write(" ".join(r) + "\n")
            cur = None
            continue
        if re.match(r"^-?[0-9]", s):
            vals = [v for v in s.split(",") if v.strip()]
            assert len(vals) == 3, s
            rows.append([v.strip() for v in vals])
    # atom tables
    pat = re.compile(r'\s*(_Charmmr|_Charmmc|_ACE)\.insert\(map<string,float>::value_type\("([^"]*)",\s*([-+0-9.eE]+)\)\);')
    out = {k: [] for k in MAP_TABLES}
    for line in text:
        m = pat.match(line)
        if m:
            out[m.group(1)].append((m.group(2), m.group(3)))
    for k, fn in MAP_TABLES.items():
        with open(os.path.join(HERE, fn), "w") as f:
            f.write("# key<TAB>value; from src/parameter.cpp %s.insert (first occurrence of a key wins)\n" % k)
            for key, val in out[k]:
                assert "\t" not in key
                f.write("%s\t%s\n" % (key, val))
        print(fn, len(out[k]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/parameter.cpp")
