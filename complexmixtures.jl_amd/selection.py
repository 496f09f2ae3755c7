"""AtomSelection, mirroring src/AtomSelection.jl:5-59 and :290-372.

Indices are 1-based positions in the full coordinate vector of the
trajectory, as in the reference.
"""
import warnings

import numpy as np

from .options import ArgumentError


def set_nmols_natomspermol(indices, nmols, natomspermol):
    """src/AtomSelection.jl:5-37"""
    natoms = len(indices)
    if natoms == 0:
        raise ArgumentError("Vector of atom indices is empty.\n")
    if nmols == 0 and natomspermol == 0:
        raise ArgumentError("\nSet nmols or natomspermol when defining a selection.\n")
    if nmols != 0:
        if natoms % nmols != 0:
            raise ArgumentError("\nNumber of atoms in selection must be a multiple of nmols.\n")
        natomspermol = natoms // nmols
    else:
        if natoms % natomspermol != 0:
            raise ArgumentError("\nNumber of atoms in selection must be a multiple of natomspermols.\n")
        nmols = natoms // natomspermol
    return nmols, natomspermol


class AtomSelection:
    """AtomSelection(indices; nmols=0, natomspermol=0, group_atom_indices=[], group_names=[])"""

    def __init__(self, indices, *, nmols=0, natomspermol=0, group_atom_indices=None, group_names=None):
        indices = [int(i) for i in indices]
        nmols, natomspermol = set_nmols_natomspermol(indices, int(nmols), int(natomspermol))
        group_atom_indices = [list(map(int, g)) for g in (group_atom_indices or [])]
        group_names = [str(n) for n in (group_names or [])]
        custom_groups = len(group_atom_indices) > 0
        if not custom_groups and group_names:
            if len(group_names) != natomspermol:
                raise ArgumentError("The length of the group_names vector does not correspond to the number of atoms "
                                    "per molecule,but no atom groups vector was provided. ")
        if custom_groups:
            index_set = set(indices)
            for inds in group_atom_indices:
                if any(inds[k] > inds[k + 1] for k in range(len(inds) - 1)):
                    warnings.warn("\nGroup indices are not sorted. The array will be sorted for faster search.\n")
                    inds.sort()
                if len(set(inds)) != len(inds):
                    raise ArgumentError("\nFound repeated indices in custom group atom indices.\n"
                                        "The group atom indices must have unique entries.\n")
                if any(i not in index_set for i in inds):
                    raise ArgumentError("\nGroup atom indices not found in the the current AtomSelection main atomic "
                                        "indices.\n")
            if not group_names:
                warnings.warn("\nVector of group atom indices was provided but vector of group names is empty. \n"
                              "The group contributions will be only retrieved by the group indices.\n")
            elif len(group_atom_indices) != len(group_names):
                raise ArgumentError("\nThe vector of group atom indices has a different number of elements than the "
                                    "vector of group names.\n")
        self.nmols = nmols
        self.natomspermol = natomspermol
        self.indices = indices
        self.custom_groups = custom_groups
        self.group_atom_indices = group_atom_indices
        self.group_names = group_names

    def __eq__(self, other):
        return (isinstance(other, AtomSelection) and self.nmols == other.nmols
                and self.natomspermol == other.natomspermol and self.indices == other.indices
                and self.custom_groups == other.custom_groups and self.group_atom_indices == other.group_atom_indices
                and self.group_names == other.group_names)

    def __repr__(self):
        ng = len(self.group_atom_indices) if self.custom_groups else self.natomspermol
        return (f"AtomSelection \n    {len(self.indices)} atoms belonging to {self.nmols} molecule(s).\n"
                f"    Atoms per molecule: {self.natomspermol}\n    Number of groups: {ng}")

    @property
    def ngroups(self):
        return len(self.group_atom_indices) if self.custom_groups else self.natomspermol

    def to_dict(self):
        return {"nmols": self.nmols, "natomspermol": self.natomspermol, "indices": list(self.indices),
                "custom_groups": self.custom_groups, "group_atom_indices": [list(g) for g in self.group_atom_indices],
                "group_names": list(self.group_names)}

    @classmethod
    def from_dict(cls, d):
        s = object.__new__(cls)
        s.nmols, s.natomspermol = int(d["nmols"]), int(d["natomspermol"])
        s.indices = [int(i) for i in d["indices"]]
        s.custom_groups = bool(d["custom_groups"])
        s.group_atom_indices = [list(map(int, g)) for g in d["group_atom_indices"]]
        s.group_names = [str(n) for n in d["group_names"]]
        return s


def natoms(sel):
    return len(sel.indices)


def group_csr(sel, atom_positions, groups_of):
    """CSR map (ptr, idx) from atoms to the custom groups of `groups_of`
    (update_group_count!, src/update_counters.jl:22-36): entry k lists the
    groups whose sorted index vector contains sel.indices[atom_positions[k]].
    Returns (None, None) without custom groups."""
    if not groups_of.custom_groups:
        return None, None
    where = {}
    for g, inds in enumerate(groups_of.group_atom_indices):
        for a in inds:
            where.setdefault(a, []).append(g)
    ptr = np.zeros(len(atom_positions) + 1, dtype=np.int32)
    idx = []
    for k, p in enumerate(atom_positions):
        gs = where.get(sel.indices[p], [])
        idx.extend(gs)
        ptr[k + 1] = ptr[k] + len(gs)
    return ptr, np.array(idx if idx else [0], dtype=np.int32)
