"""Writes the force-field fixtures from the reference's example inputs (data
only: coordinates and PDB records, no code).

  forcefields/waterbox_996.npz   the 996-water liquid box of
      examples/waterbox/waterbox_31ang.pdb: positions in units of 0.001 A
      (int32, exactly the PDB's three decimals), atom order O H1 H2 per
      water, box edge in A.
  forcefields/ethane_cluster.pdb  examples/ethane_water_charge_only/
      solvated_ethane_from_openmm_setup.pdb cut to the ethane molecule
      (with its CONECT records) and the waters whose O lies within 0.7 nm
      of the ethane's centre.

The force-field XML files beside them (swm4.xml, swm6.xml,
ethane_water.xml) are the reference's parameter files
(examples/parameters/, examples/ethane_water_charge_only/), copied as they
are.  Run in a tree that has the reference at /root/reference.
"""
import os

import numpy as np

REF = "/root/reference/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "forcefields")


def atom_lines(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    return lines, [ln for ln in lines if ln.startswith(("ATOM", "HETATM"))]


def waterbox():
    lines, atoms = atom_lines(os.path.join(REF, "waterbox", "waterbox_31ang.pdb"))
    names = [ln[12:16].strip() for ln in atoms]
    assert names == ["O", "H1", "H2"] * (len(atoms) // 3)
    mA = np.array([[round(float(ln[k:k + 8]) * 1000.0) for k in (30, 38, 46)] for ln in atoms], dtype=np.int32)
    box = [ln for ln in lines if ln.startswith("CRYST1")][0]
    edge = float(box[6:15])
    np.savez_compressed(os.path.join(OUT, "waterbox_996.npz"), positions_mA=mA, box_A=np.float64(edge))
    print("waterbox_996.npz:", mA.shape, edge)


def ethane_cluster(radius_nm=0.7):
    lines, atoms = atom_lines(os.path.join(REF, "ethane_water_charge_only", "solvated_ethane_from_openmm_setup.pdb"))
    eth = [ln for ln in atoms if ln[17:20].strip() != "HOH"]
    xyz = lambda ln: np.array([float(ln[k:k + 8]) for k in (30, 38, 46)]) / 10.0
    centre = np.mean([xyz(ln) for ln in eth], axis=0)
    waters = [ln for ln in atoms if ln[17:20].strip() == "HOH"]
    keep = []
    for w in range(0, len(waters), 3):
        o = waters[w]
        assert o[12:16].strip() == "O"
        if np.linalg.norm(xyz(o) - centre) < radius_nm:
            keep.extend(waters[w:w + 3])
    conect = [ln for ln in lines if ln.startswith("CONECT")]
    out = [ln for ln in lines if ln.startswith("CRYST1")] + eth + ["TER"] + keep + conect + ["END"]
    with open(os.path.join(OUT, "ethane_cluster.pdb"), "w") as fh:
        fh.write("\n".join(out) + "\n")
    print("ethane_cluster.pdb:", len(eth), "ethane atoms,", len(keep) // 3, "waters")


if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    waterbox()
    ethane_cluster()
