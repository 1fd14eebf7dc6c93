"""Copies the lattice databases that the reference's own simulation scripts
load (pytest/sim-lbeg.py: dbs = ['tbtd.db', 'dtl.db', 'trst.db', 'ccl.db'])
into this folder as fixtures: the low-energy transport (tbtd.db, with
bunchers and dipoles), the DTL (dtl.db), the transition region (trst.db)
and the CCL (ccl.db).  Data only (SQLite tables, views and triggers); run
once where the reference tree is present:

    python tests/golden/make_lattice_fixtures.py /root/reference/db
"""
import os
import shutil
import sqlite3
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    for name in ("tbtd.db", "dtl.db", "trst.db", "ccl.db"):
        dst = os.path.join(HERE, name)
        shutil.copyfile(os.path.join(src, name), dst)
        os.chmod(dst, 0o644)
        # integrity check of the copy
        con = sqlite3.connect(f"file:{dst}?mode=ro", uri=True)
        n = con.execute("select count(*) from linac").fetchone()[0]
        con.close()
        print(f"{name}: {n} linac rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/db")
