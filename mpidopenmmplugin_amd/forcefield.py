"""Force-field XML to MPIDForce: the reference's ``MPIDGenerator``
(python/mpidplugin.i:534-1052) with the small part of OpenMM's
``app.ForceField`` / ``app.Topology`` / ``app.PDBFile`` it relies on.

OpenMM's Python application layer is not part of this package, so this
module carries a minimal stand-in for it:

  * ``Topology`` (atoms with name and element, residues, bonds) and
    ``PDBFile`` (ATOM/HETATM, CONECT and CRYST1 records; positions in nm);
  * ``ForceField``: reads ``<AtomTypes>``, ``<Residues>`` (templates) and
    ``<MPIDForce>`` from one or more XML files, ignores the other force
    sections, and matches every residue of a topology to a template by its
    bond graph and elements (OpenMM matches templates the same way, by graph
    and not by atom names);
  * ``createSystem`` returns a ``System`` that holds the particle masses and
    the one ``MPIDForce``.

Lengths are plain floats in nm (the reference passes OpenMM unit
quantities: ``nonbondedCutoff=8*angstrom`` is 0.8 here).

What follows the reference exactly (each function cites its lines): the axis
type derived from the signs of kz/kx/ky, the multipole and polarisability
records, the four passes that pick the frame-defining atoms, the 1-2/1-3/1-4
covalent maps and every ``createSystem`` argument.  The frame atoms depend on
the iteration order of Python sets of atom indices in the reference; the same
sets are built here in the same way, so the same atoms are picked.
"""
from __future__ import annotations

import warnings
import xml.etree.ElementTree as ET

from .mpidforce import MPIDForce

# nonbonded methods of openmm.app.forcefield that MPIDGenerator.createForce accepts
NoCutoff = "NoCutoff"
CutoffNonPeriodic = "CutoffNonPeriodic"
CutoffPeriodic = "CutoffPeriodic"
Ewald = "Ewald"
PME = "PME"
LJPME = "LJPME"
_METHODS = {NoCutoff: MPIDForce.NoCutoff, PME: MPIDForce.PME, LJPME: MPIDForce.PME}  # mpidplugin.i:732-734


# ---------------------------------------------------------------------------
# topology
# ---------------------------------------------------------------------------
class Atom:
    def __init__(self, index, name, element, residue):
        self.index = index
        self.name = name
        self.element = element
        self.residue = residue


class Residue:
    def __init__(self, index, name, chain="A", resid=None):
        self.index = index
        self.name = name
        self.chain = chain
        self.id = resid
        self.atoms = []


class Topology:
    """Atoms grouped in residues, and bonds between atom indices."""

    def __init__(self):
        self.atoms = []
        self.residues = []
        self.bonds = []
        self.box = None  # edge lengths in nm, or None

    def addResidue(self, name, chain="A", resid=None):
        r = Residue(len(self.residues), name, chain, resid)
        self.residues.append(r)
        return r

    def addAtom(self, name, element, residue):
        a = Atom(len(self.atoms), name, element, residue)
        self.atoms.append(a)
        residue.atoms.append(a)
        return a

    def addBond(self, a, b):
        i = a.index if isinstance(a, Atom) else int(a)
        j = b.index if isinstance(b, Atom) else int(b)
        if i == j:
            raise ValueError("an atom cannot be bonded to itself")
        self.bonds.append((i, j))

    def getNumAtoms(self):
        return len(self.atoms)

    def getPeriodicBoxVectors(self):
        return self.box


def water_topology(nwaters, order=("O", "H1", "H2"), resname="HOH"):
    """nwaters water residues, atoms in the given name order, O-H bonds."""
    top = Topology()
    for _ in range(nwaters):
        r = top.addResidue(resname)
        atoms = {name: top.addAtom(name, name[0], r) for name in order}
        top.addBond(atoms["O"], atoms["H1"])
        top.addBond(atoms["O"], atoms["H2"])
    return top


class PDBFile:
    """The records of a PDB file that a force-field assignment needs: atoms
    (first MODEL only), residues, CONECT bonds and the CRYST1 box.  Bonds
    inside residues that carry no CONECT records are added from the
    matching force-field template in ``ForceField.createSystem``."""

    def __init__(self, path_or_lines):
        lines = path_or_lines
        if isinstance(path_or_lines, str):
            with open(path_or_lines) as fh:
                lines = fh.read().splitlines()
        self.topology = Topology()
        self.positions = []
        serial_to_index = {}
        last_key = None
        res = None
        for line in lines:
            rec = line[:6].strip()
            if rec == "ENDMDL":
                break
            if rec == "CRYST1":
                self.topology.box = [float(line[6:15]) / 10.0, float(line[15:24]) / 10.0, float(line[24:33]) / 10.0]
            elif rec in ("ATOM", "HETATM"):
                name = line[12:16].strip()
                resname = line[17:21].strip()
                chain = line[21:22]
                resid = line[22:27].strip()
                key = (chain, resid, resname)
                if key != last_key or res is None:
                    res = self.topology.addResidue(resname, chain, resid)
                    last_key = key
                element = line[76:78].strip() if len(line) >= 78 else ""
                if not element:
                    element = "".join(c for c in name if c.isalpha())[:1] or "X"
                element = element[0].upper() + element[1:].lower()
                a = self.topology.addAtom(name, element, res)
                serial = line[6:11].strip()
                if serial:
                    serial_to_index[serial] = a.index
                self.positions.append([float(line[30:38]) / 10.0, float(line[38:46]) / 10.0,
                                       float(line[46:54]) / 10.0])
            elif rec == "TER":
                last_key = None
        seen = set()
        for line in lines:
            if line[:6].strip() != "CONECT":
                continue
            fields = [line[k:k + 5].strip() for k in range(6, len(line), 5)]
            fields = [f for f in fields if f]
            if not fields or fields[0] not in serial_to_index:
                continue
            i = serial_to_index[fields[0]]
            for f in fields[1:]:
                if f in serial_to_index:
                    j = serial_to_index[f]
                    if (min(i, j), max(i, j)) not in seen and i != j:
                        seen.add((min(i, j), max(i, j)))
                        self.topology.addBond(i, j)

    def getTopology(self):
        return self.topology

    def getPositions(self):
        return self.positions


# ---------------------------------------------------------------------------
# the generator
# ---------------------------------------------------------------------------
def set_axis_type(kIndices):
    """MPIDGenerator.setAxisType (mpidplugin.i:553-601): the axis type from
    which of kz, kx, ky are given and which carry a leading '-'.  Strips the
    '-' in place and pads kIndices to [type, kz, kx, ky]."""
    n = len(kIndices)
    ky = kx = kz = ""
    ky_neg = kx_neg = kz_neg = False
    if n > 3:
        ky = kIndices[3]
        if ky.startswith("-"):
            ky = kIndices[3] = ky[1:]
            ky_neg = True
    if n > 2:
        kx = kIndices[2]
        if kx.startswith("-"):
            kx = kIndices[2] = kx[1:]
            kx_neg = True
    if n > 1:
        kz = kIndices[1]
        if kz.startswith("-"):
            kz = kIndices[1] = kz[1:]
            kz_neg = True
    while len(kIndices) < 4:
        kIndices.append("")
    axis = MPIDForce.ZThenX
    if not kz:
        axis = MPIDForce.NoAxisType
    if kz and not kx:
        axis = MPIDForce.ZOnly
    if (kz and kz_neg) or (kx and kx_neg):
        axis = MPIDForce.Bisector
    if kx and kx_neg and ky and ky_neg:
        axis = MPIDForce.ZBisect
    if kz and kz_neg and kx and kx_neg and ky and ky_neg:
        axis = MPIDForce.ThreeFold
    return axis


_QUAD = ("qXX", "qXY", "qYY", "qXZ", "qYZ", "qZZ")
_OCT = ("oXXX", "oXXY", "oXYY", "oYYY", "oXXZ", "oXYZ", "oYYZ", "oXZZ", "oYZZ", "oZZZ")


class MPIDGenerator:
    """Per-type multipole records of one or more ``<MPIDForce>`` elements
    and the MPIDForce they give for a topology (mpidplugin.i:534-1052)."""

    def __init__(self, forceField, scaleFactor14, defaultTholeWidth):
        self.forceField = forceField
        self.scaleFactor14 = scaleFactor14
        self.defaultTholeWidth = defaultTholeWidth
        self.typeMap = {}

    @staticmethod
    def parseElement(element, forceField):
        """MPIDGenerator.parseElement, mpidplugin.i:606-726."""
        existing = [g for g in forceField._generators if isinstance(g, MPIDGenerator)]
        if not existing:
            gen = MPIDGenerator(forceField, element.get("coulomb14scale", None), element.get("defaultTholeWidth", None))
            forceField._generators.append(gen)
        else:
            gen = existing[0]
            if gen.scaleFactor14 != element.get("coulomb14scale", None):
                raise ValueError("Found multiple MPIDForce tags with different coulomb14scale arguments")
            if gen.defaultTholeWidth != element.get("defaultTholeWidth", None):
                raise ValueError("Found multiple MPIDForce tags with different defaultTholeWidth arguments")
        for atom in element.findall("Multipole"):
            types = forceField._findAtomTypes(atom.attrib)
            if types is None or "type" not in atom.attrib:
                raise ValueError("MPIDGenerator: error getting type for multipole: %s" %
                                 atom.attrib.get("class", atom.attrib.get("type")))
            kIndices = [atom.attrib["type"]]
            for k in ("kz", "kx", "ky"):
                if atom.attrib.get(k):
                    kIndices.append(atom.attrib[k])
            axis = set_axis_type(kIndices)
            rec = {
                "classIndex": atom.attrib["type"],
                "kIndices": kIndices,
                "charge": float(atom.get("c0")),
                "dipole": [float(atom.get(k, 0.0)) for k in ("dX", "dY", "dZ")],
                "quadrupole": [float(atom.get(k, 0.0)) for k in _QUAD],
                "octopole": [float(atom.get(k, 0.0)) for k in _OCT],
                "axisType": axis,
            }
            for t in types:
                gen.typeMap.setdefault(t, []).append(dict(rec))
        for atom in element.findall("Polarize"):
            types = forceField._findAtomTypes(atom.attrib)
            if types is None:
                raise ValueError("MPIDGenerator: error getting type for polarize: %s" %
                                 atom.attrib.get("class", atom.attrib.get("type")))
            classIndex = atom.attrib["type"]
            pol = [float(atom.attrib[k]) for k in ("polarizabilityXX", "polarizabilityYY", "polarizabilityZZ")]
            thole = float(atom.attrib["thole"])
            for t in types:
                if t not in gen.typeMap:
                    raise ValueError("MPIDGenerator: polarize type not present: %s" % atom.attrib["type"])
                hit = False
                for rec in gen.typeMap[t]:
                    if rec["classIndex"] == classIndex:
                        rec["polarizability"] = pol
                        rec["thole"] = thole
                        hit = True
                if not hit:
                    raise ValueError("MPIDGenerator: error getting type for polarize: class index=%s not in "
                                     "multipole list?" % atom.attrib.get("class", classIndex))

    @staticmethod
    def _bonded_sets(natoms, bonds):
        """1-2, 1-3 and 1-4 partner sets (mpidplugin.i:800-846; the 1-2 sets
        as openmm.app's AmoebaVdwGenerator.getBondedParticleSets builds them,
        one ``add`` per bond end in bond order)."""
        b12 = [set() for _ in range(natoms)]
        for i, j in bonds:
            b12[i].add(j)
            b12[j].add(i)
        b13 = []
        for i in range(natoms):
            s = set()
            for j in b12[i]:
                s = s.union(b12[j])
            s = s - b12[i] - {i}
            b13.append(set(sorted(s)))
        b14 = []
        for i in range(natoms):
            s = set()
            for j in b13[i]:
                s = s.union(b12[j])
            s = s - b12[i] - b13[i] - {i}
            b14.append(set(sorted(s)))
        return b12, b13, b14

    @staticmethod
    def _frame(atomIndex, multipoleList, types, b12, b13):
        """The multipole record and frame atoms of one atom: the four passes
        of mpidplugin.i:850-1021 (1-2 partners; a 1-2 z atom with 1-3 x/y
        atoms bonded to it; a z atom only; no frame atoms).  Returns
        (record, z, x, y) or None."""
        bonded = b12[atomIndex]
        # pass 1: z, x (and y) among the 1-2 partners
        for rec in multipoleList:
            _, kz, kx, ky = rec["kIndices"]
            for zi in bonded:
                if types[zi] != kz:
                    continue
                for xi in bonded:
                    if xi == zi or types[xi] != kx:
                        continue
                    if not ky:
                        z, x = zi, xi
                        if types[xi] == types[zi] and x < z:
                            z, x = x, z
                        else:
                            for x1 in bonded:
                                if types[x1] == kx and x1 != z and x1 < x:
                                    x = x1
                        return rec, z, x, -1
                    for yi in bonded:
                        if yi == zi or yi == xi or types[yi] != ky:
                            continue
                        return rec, zi, xi, yi
        # pass 2: z a 1-2 partner, x (and y) 1-3 partners bonded to z
        b13i = b13[atomIndex]
        for rec in multipoleList:
            _, kz, kx, ky = rec["kIndices"]
            for zi in bonded:
                if types[zi] != kz:
                    continue
                for xi in b13i:
                    if xi == zi or types[xi] != kx or zi not in b12[xi]:
                        continue
                    if not ky:
                        x = xi
                        for x1 in b13i:
                            if types[x1] == kx and x1 != zi and zi in b12[x1] and x1 < x:
                                x = x1
                        return rec, zi, x, -1
                    for yi in b13i:
                        if yi == zi or yi == xi or types[yi] != ky or zi not in b12[yi]:
                            continue
                        return rec, zi, xi, yi
        # pass 3: a z atom only
        for rec in multipoleList:
            _, kz, kx, _ = rec["kIndices"]
            for zi in bonded:
                if not kx and kz == types[zi]:
                    return rec, zi, -1, -1
        # pass 4: no frame atoms
        for rec in multipoleList:
            if not rec["kIndices"][1]:
                return rec, -1, -1, -1
        return None

    def createForce(self, system, data, nonbondedMethod, nonbondedCutoff, args):
        """MPIDGenerator.createForce, mpidplugin.i:730-1050."""
        force = MPIDForce()
        system.addForce(force)
        if nonbondedMethod not in _METHODS:
            raise ValueError("MPIDForce: input cutoff method not available.")
        force.setNonbondedMethod(_METHODS[nonbondedMethod])
        force.setCutoffDistance(float(nonbondedCutoff))
        if "ewaldErrorTolerance" in args:
            force.setEwaldErrorTolerance(float(args["ewaldErrorTolerance"]))
        force.setPolarizationType(MPIDForce.Extrapolated)
        if "polarization" in args:
            p = str(args["polarization"]).lower()
            if p == "direct":
                force.setPolarizationType(MPIDForce.Direct)
            elif p == "mutual":
                force.setPolarizationType(MPIDForce.Mutual)
            elif p == "extrapolated":
                force.setPolarizationType(MPIDForce.Extrapolated)
            else:
                raise ValueError("MPIDForce: invalide polarization type: " + str(args["polarization"]))
        for key, setter in (("coulomb14scale", force.set14ScaleFactor),
                            ("defaultTholeWidth", force.setDefaultTholeWidth)):
            argval = float(args[key]) if key in args else None
            mine = self.scaleFactor14 if key == "coulomb14scale" else self.defaultTholeWidth
            myval = float(mine) if mine else None
            if argval is not None:
                if myval is not None and myval != argval:
                    warnings.warn("Conflicting {} values found in forcefield file ({}) and createSystem args ({}).  "
                                  "Using the value from createSystem's arguments".format(key, myval, argval))
                setter(argval)
            elif myval is not None:
                setter(myval)
        if "aEwald" in args:
            force.setAEwald(float(args["aEwald"]))
        if "pmeGridDimensions" in args:
            force.setPmeGridDimensions(args["pmeGridDimensions"])
        if "mutualInducedMaxIterations" in args:
            force.setMutualInducedMaxIterations(int(args["mutualInducedMaxIterations"]))
        if "mutualInducedTargetEpsilon" in args:
            force.setMutualInducedTargetEpsilon(float(args["mutualInducedTargetEpsilon"]))

        types = data.atomType
        b12, b13, b14 = self._bonded_sets(len(types), data.bonds)
        for i, atom in enumerate(data.atoms):
            t = types[i]
            if t not in self.typeMap:
                raise ValueError("No multipole type for atom %s %s %d" % (atom.name, atom.residue.name,
                                                                          atom.residue.index))
            found = self._frame(i, self.typeMap[t], types, b12, b13)
            if found is None:
                raise ValueError("Atom %s of %s %d was not assigned." % (atom.name, atom.residue.name,
                                                                         atom.residue.index))
            rec, z, x, y = found
            idx = force.addMultipole(rec["charge"], rec["dipole"], rec["quadrupole"], rec["octopole"],
                                     rec["axisType"], z, x, y, rec.get("thole", 0.0),
                                     rec.get("polarizability", [0.0, 0.0, 0.0]))
            if idx != i:
                raise ValueError("Atom %s of %s %d is out of sync!." % (atom.name, atom.residue.name,
                                                                        atom.residue.index))
            # in set iteration order, as the reference passes tuple(set)
            force.setCovalentMap(i, MPIDForce.Covalent12, list(b12[i]))
            force.setCovalentMap(i, MPIDForce.Covalent13, list(b13[i]))
            force.setCovalentMap(i, MPIDForce.Covalent14, list(b14[i]))
        return force


# ---------------------------------------------------------------------------
# force field
# ---------------------------------------------------------------------------
class _Template:
    def __init__(self, name):
        self.name = name
        self.atoms = []  # (name, type)
        self.bonds = []  # (i, j) template atom indices


class _Data:
    """What createForce reads of openmm.app's ForceField._SystemData."""

    def __init__(self, atoms, atomType, bonds):
        self.atoms = atoms
        self.atomType = atomType
        self.bonds = bonds


class System:
    """Particle masses and forces: the part of openmm.System that
    ``createSystem`` fills here."""

    def __init__(self):
        self._masses = []
        self._forces = []
        self.box = None

    def addParticle(self, mass):
        self._masses.append(float(mass))
        return len(self._masses) - 1

    def getNumParticles(self):
        return len(self._masses)

    def getParticleMass(self, i):
        return self._masses[i]

    def addForce(self, force):
        self._forces.append(force)
        return len(self._forces) - 1

    def getNumForces(self):
        return len(self._forces)

    def getForce(self, i):
        return self._forces[i]

    def getForces(self):
        return list(self._forces)


class ForceField:
    """Atom types, residue templates and ``<MPIDForce>`` records of one or
    more force-field XML files (other force sections are not read)."""

    def __init__(self, *files):
        self._types = {}  # name -> (class, element, mass)
        self._templates = {}
        self._generators = []
        self.loadFile(*files)

    def loadFile(self, *files):
        """Reads the atom types and residue templates of every file before
        any <MPIDForce> section, as openmm.app.ForceField does, so that a
        force section may name types defined in another file."""
        roots = [f if isinstance(f, ET.Element) else ET.parse(f).getroot() for f in files]
        for root in roots:
            self._load_types(root)
        for root in roots:
            for el in root.findall("MPIDForce"):
                MPIDGenerator.parseElement(el, self)

    def _load_types(self, root):
        for t in root.findall("AtomTypes/Type"):
            self._types[t.get("name")] = (t.get("class"), t.get("element"), float(t.get("mass", "0")))
        for r in root.findall("Residues/Residue"):
            tpl = _Template(r.get("name"))
            names = {}
            for a in r.findall("Atom"):
                names[a.get("name")] = len(tpl.atoms)
                tpl.atoms.append((a.get("name"), a.get("type")))
            for b in r.findall("Bond"):
                if b.get("from") is not None:
                    tpl.bonds.append((int(b.get("from")), int(b.get("to"))))
                else:
                    tpl.bonds.append((names[b.get("atomName1")], names[b.get("atomName2")]))
            self._templates[tpl.name] = tpl

    def _findAtomTypes(self, attrib):
        """The types an entry names: its ``type``, or every type of its
        ``class`` (openmm.app.ForceField._findAtomTypes with one atom);
        None when it names neither an existing type nor class."""
        if "type" in attrib:
            return [attrib["type"]] if attrib["type"] in self._types else None
        if "class" in attrib:
            ts = [n for n, (c, _, _) in self._types.items() if c == attrib["class"]]
            return ts or None
        return None

    def _element(self, type_name):
        return self._types[type_name][1]

    def _match(self, residue, internal):
        """Template atom index for each residue atom: a template with the
        same elements whose bond graph is the residue's, matched by a
        depth-first search (residue bonds are taken from CONECT records;
        a residue that has none takes the template's bonds once its atoms
        are matched by element and, for the first atom of each element, by
        name)."""
        n = len(residue.atoms)
        local = {a.index: k for k, a in enumerate(residue.atoms)}
        adj = [set() for _ in range(n)]
        for i, j in internal:
            adj[local[i]].add(local[j])
            adj[local[j]].add(local[i])
        elements = [a.element for a in residue.atoms]
        candidates = [t for t in self._templates.values() if len(t.atoms) == n]
        # the template of the same name first
        candidates.sort(key=lambda t: t.name != residue.name)
        for tpl in candidates:
            telem = [self._element(ty) for _, ty in tpl.atoms]
            if sorted(telem) != sorted(elements):
                continue
            tadj = [set() for _ in range(n)]
            for i, j in tpl.bonds:
                tadj[i].add(j)
                tadj[j].add(i)
            if not internal:
                # no bonds given: match by name where names agree, else by element order
                by_name = {nm: k for k, (nm, _) in enumerate(tpl.atoms)}
                m, used = [None] * n, set()
                for k, a in enumerate(residue.atoms):
                    t = by_name.get(a.name)
                    if t is not None and t not in used and telem[t] == elements[k]:
                        m[k] = t
                        used.add(t)
                for k in range(n):
                    if m[k] is None:
                        t = next((t for t in range(n) if t not in used and telem[t] == elements[k]), None)
                        if t is None:
                            break
                        m[k] = t
                        used.add(t)
                if None not in m:
                    return tpl, m, True
                continue
            if sorted(len(s) for s in tadj) != sorted(len(s) for s in adj):
                continue
            m = [None] * n
            used = [False] * n

            def dfs(k):
                if k == n:
                    return True
                for t in range(n):
                    if used[t] or telem[t] != elements[k] or len(tadj[t]) != len(adj[k]):
                        continue
                    if any(m[q] is not None and (m[q] in tadj[t]) != (q in adj[k]) for q in range(n)):
                        continue
                    m[k] = t
                    used[t] = True
                    if dfs(k + 1):
                        return True
                    m[k] = None
                    used[t] = False
                return False

            if dfs(0):
                return tpl, m, False
        raise ValueError("No template found for residue %d (%s).  The set of atoms matches no residue "
                         "template of the force field." % (residue.index, residue.name))

    def createSystem(self, topology, nonbondedMethod=NoCutoff, nonbondedCutoff=1.0, **args):
        """A ``System`` with the masses of the typed atoms and the MPIDForce
        of the ``<MPIDForce>`` records (MPIDGenerator.createForce)."""
        if not self._generators:
            raise ValueError("the force field has no <MPIDForce> records")
        n = topology.getNumAtoms()
        atom_type = [None] * n
        bonds = list(topology.bonds)
        by_res = {}
        for i, j in topology.bonds:
            ri, rj = topology.atoms[i].residue.index, topology.atoms[j].residue.index
            if ri == rj:
                by_res.setdefault(ri, []).append((i, j))
        for res in topology.residues:
            internal = by_res.get(res.index, [])
            tpl, m, add_bonds = self._match(res, internal)
            for k, a in enumerate(res.atoms):
                atom_type[a.index] = tpl.atoms[m[k]][1]
            if add_bonds:
                inv = {m[k]: res.atoms[k].index for k in range(len(res.atoms))}
                for i, j in tpl.bonds:
                    bonds.append((inv[i], inv[j]))
        topology.bonds = bonds
        system = System()
        for t in atom_type:
            system.addParticle(self._types[t][2])
        data = _Data(topology.atoms, atom_type, bonds)
        self._generators[0].createForce(system, data, nonbondedMethod, nonbondedCutoff, args)
        system.box = topology.box
        return system
