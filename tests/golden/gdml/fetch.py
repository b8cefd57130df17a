"""Copies the reference's GDML test geometries used by tests/test_gdml.py
(persistency/gdml/gdmls/*.gdml, data files of the reference's GDML tests)
into this folder.  Run in a checkout that has the reference at
/root/reference; the copies are committed."""
import os
import shutil

SRC = "/root/reference/persistency/gdml/gdmls"
FILES = ["aPosition", "aRotation", "auxiliary", "oneBox", "oneConeSegment", "oneGeneralTrapezoid", "onePolycone",
         "onePolyhedron", "oneTesselatedSolid", "oneTrapezoid", "oneTube", "twoBoxes",
         # refused: a solid outside the implemented families, a scaled physvol
         "oneOrb", "scaledPhysvol", "aConstant", "oneCutTube", "solidSubtraction",
         "solidIntersection", "solidUnion", "oneSphere", "oneTorus",
         # materials (read and ignored), a rotated daughter in a boolean world
         "compositeMaterial", "mixtureElement", "mixtureMaterial", "simpleMaterial", "unionWIthDetector"]

if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        shutil.copy(os.path.join(SRC, f + ".gdml"), os.path.join(here, f + ".gdml"))
    print(f"copied {len(FILES)} files")
