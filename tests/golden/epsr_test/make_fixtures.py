"""Copies the reference's own EPSR test inputs into this folder as fixtures (gzip, mtime 0).

Sources (data files of the reference, not code):
  EPSR-test/200DMSO400Water.ato, .pcof, .EPSR.inp  -- the 200 DMSO + 400 water test box
      (3200 atoms, cubic, L = 33.1213 A) with its saved ran1 state and its 116-term EP
  EPSR-test/DMSO.ato, test_problems/200DMSO400Water/EPSR-test-ex46/WaterMolecule.ato
      -- single-molecule files
  test_problems/200DMSO400Water/EPSR-test-ex8/DMSODWater{D,Null,H}.NWTStot.wts
      -- the neutron weights of the test box's three datasets
Run from the repository root where /root/reference exists:
  python tests/golden/epsr_test/make_fixtures.py
"""
import gzip
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["EPSR-test/200DMSO400Water.ato", "EPSR-test/200DMSO400Water.pcof", "EPSR-test/200DMSO400Water.EPSR.inp",
         "EPSR-test/DMSO.ato", "test_problems/200DMSO400Water/EPSR-test-ex46/WaterMolecule.ato",
         "test_problems/200DMSO400Water/EPSR-test-ex8/DMSODWaterD.NWTStot.wts",
         "test_problems/200DMSO400Water/EPSR-test-ex8/DMSODWaterNull.NWTStot.wts",
         "test_problems/200DMSO400Water/EPSR-test-ex8/DMSODWaterH.NWTStot.wts"]

for rel in FILES:
    dst = os.path.join(HERE, os.path.basename(rel) + ".gz")
    with open(os.path.join(REF, rel), "rb") as f, open(dst, "wb") as raw:
        with gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0) as g:
            shutil.copyfileobj(f, g)
    print(dst, os.path.getsize(dst))
