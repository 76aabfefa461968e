"""Phylogenetic tree conventions of the reference (host side).

* Newick parsing: src/Tree.jl:114-171 (``treefromnewickhelper``/``gettreefromnewick``).
* Rooting: src/Tree.jl:45-55 (``roottree``): child ``index`` of the input root is
  spliced under a new root, and its branch length is halved between it and the
  old root.  The input must therefore have a trifurcating root for the result
  to be binary.
* Node numbering: src/Tree.jl:190-208 (``annotatetree``/``getnodelist``): pre-order,
  root first.  This module uses 0-based indices (reference index - 1).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np


@dataclass
class TreeNode:
    branchlength: float = 0.0
    name: str = ""
    children: List["TreeNode"] = field(default_factory=list)
    parent: Optional["TreeNode"] = None
    nodeindex: int = -1
    seqindex: int = -1

    def isleaf(self) -> bool:
        return len(self.children) == 0


def _split_top(s: str) -> List[str]:
    out, depth, cur = [], 0, []
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    return out


def gettreefromnewick(newick: str) -> TreeNode:
    """Parse a Newick string, following src/Tree.jl:114-171."""
    s = newick.strip().rstrip(";").strip()

    def parse(t: str) -> TreeNode:
        t = t.strip()
        close = t.rfind(")")
        if close >= 0:
            tag = t[close + 1:]
            inner = t[1:close]
            kids = _split_top(inner)
        else:
            tag, kids = t, []
        parts = tag.split(":")
        name = parts[0].strip() if parts else ""
        bl = float(parts[1]) if len(parts) > 1 and parts[1].strip() else 0.0
        node = TreeNode(branchlength=bl, name=name)
        for k in kids:
            c = parse(k)
            c.parent = node
            node.children.append(c)
        return node

    return parse(s)


def roottree(root: TreeNode, index: int = 0) -> TreeNode:
    """src/Tree.jl:45-55 (0-based ``index``)."""
    newroot = TreeNode(0.0, "")
    child = root.children.pop(index)
    child.parent = newroot
    newroot.children.append(child)
    br = child.branchlength / 2.0
    child.branchlength = br
    root.branchlength = br
    root.parent = newroot
    newroot.children.append(root)
    return newroot


def getnodelist(root: TreeNode) -> List[TreeNode]:
    out: List[TreeNode] = []
    stack = [root]
    while stack:
        n = stack.pop()
        out.append(n)
        stack.extend(reversed(n.children))
    return out


def annotatetree(root: TreeNode, seqnametoindex: Dict[str, int]) -> List[TreeNode]:
    """Pre-order node indices (root = 0) and leaf sequence rows, src/Tree.jl:190-199."""
    nodes = getnodelist(root)
    for i, n in enumerate(nodes):
        n.nodeindex = i
        n.seqindex = seqnametoindex.get(n.name, -1)
    return nodes


def getnewick(node: TreeNode) -> str:
    def h(n: TreeNode) -> str:
        if n.isleaf():
            return f"{n.name}:{n.branchlength!r}"
        return "(" + ",".join(h(c) for c in n.children) + f"){n.name}:{n.branchlength!r}"
    return h(node) + ";"


@dataclass
class TreeArrays:
    """Flat pre-order arrays describing a rooted binary tree (root = 0)."""
    children: np.ndarray      # (nnodes, 2) int32, -1 for leaves
    parent: np.ndarray        # (nnodes,) int32, -1 for the root
    seqindex: np.ndarray      # (nnodes,) int32, leaf -> alignment row, -1 for internal nodes
    branchlengths: np.ndarray  # (nnodes,) float64, root entry 0
    names: List[str]

    @property
    def nnodes(self) -> int:
        return len(self.parent)

    def leaves(self) -> np.ndarray:
        return np.nonzero(self.children[:, 0] < 0)[0]

    def internal(self) -> np.ndarray:
        return np.nonzero(self.children[:, 0] >= 0)[0]


def tree_arrays(root: TreeNode, seqnametoindex: Dict[str, int]) -> TreeArrays:
    nodes = annotatetree(root, seqnametoindex)
    n = len(nodes)
    ch = -np.ones((n, 2), dtype=np.int32)
    par = -np.ones(n, dtype=np.int32)
    seq = -np.ones(n, dtype=np.int32)
    bl = np.zeros(n)
    for nd in nodes:
        i = nd.nodeindex
        bl[i] = nd.branchlength
        if nd.isleaf():
            if nd.seqindex < 0:
                raise ValueError(f"leaf '{nd.name}' has no sequence in the alignment")
            seq[i] = nd.seqindex
        else:
            if len(nd.children) < 2:
                raise ValueError("internal node with fewer than 2 children: the Newick root must be "
                                 "trifurcating (src/Tree.jl:45-55 splices one child under a new root)")
            # Only children[1] and children[2] are used by the reference (src/FelsensteinCUDA.jl:340-341).
            ch[i, 0] = nd.children[0].nodeindex
            ch[i, 1] = nd.children[1].nodeindex
        for c in nd.children:
            par[c.nodeindex] = i
    return TreeArrays(ch, par, seq, bl, [nd.name for nd in nodes])
