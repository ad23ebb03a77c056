"""taxo.k2d writer: NCBI nodes.dmp/names.dmp -> the pruned, BFS-numbered
Kraken 2 taxonomy (src/taxonomy.cc:19-185, 275-287).

Only nodes that are ancestors of a mapped taxon (plus the root) are kept;
internal ids are assigned breadth first from the root with children taken in
ascending external-id order, so a parent's id is below its children's and
every node's children are contiguous.
"""
from __future__ import annotations

import struct
from collections import deque


def parse_dmp(nodes_path: str, names_path: str):
    """Returns (parent, rank, name) dicts keyed by external taxid."""
    parent, rank, name = {}, {}, {}
    with open(nodes_path) as f:
        for line in f:
            line = line.rstrip("\n")
            line = line[:-2] if len(line) >= 2 else line  # trailing "\t|"
            fields = line.split("\t|\t", 3)
            if len(fields) < 3:
                continue
            node, par = int(fields[0]), int(fields[1])
            if node == 0:
                raise ValueError("attempt to create taxonomy w/ node ID == 0")
            parent[node] = 0 if node == 1 else par
            rank[node] = fields[2]
    with open(names_path) as f:
        for line in f:
            line = line.rstrip("\n")
            line = line[:-2] if len(line) >= 2 else line
            fields = line.split("\t|\t", 4)
            if len(fields) < 4:
                continue
            if fields[3] == "scientific name":
                name[int(fields[0])] = fields[1]
    return parent, rank, name


def build_k2_taxonomy(parent: dict, rank: dict, name: dict, mapped_taxids) -> bytes:
    """Serialized taxo.k2d for the tree restricted to ancestors of mapped_taxids."""
    marked = {1}
    for t in mapped_taxids:
        while t not in marked:
            marked.add(t)
            t = parent.get(t, 0)
    children: dict[int, list[int]] = {}
    for node, par in parent.items():
        children.setdefault(par, []).append(node)
    for v in children.values():
        v.sort()
    known_ranks = sorted(set(rank.values()), key=lambda s: s.encode())
    rank_off, rank_data = {}, bytearray()
    for r in known_ranks:
        rank_off[r] = len(rank_data)
        rank_data += r.encode() + b"\0"
    nodes = [(0, 0, 0, 0, 0, 0, 0)]
    name_data = bytearray()
    ext_to_int = {0: 0, 1: 1}
    q = deque([1])
    next_id = 0
    while q:
        ext = q.popleft()
        next_id += 1
        ext_to_int[ext] = next_id
        first_child = next_id + 1 + len(q)
        count = 0
        for c in children.get(ext, []):
            if c in marked:
                q.append(c)
                count += 1
        par_int = ext_to_int.get(parent.get(ext, 0), 0)
        nodes.append((par_int, first_child, count, len(name_data), rank_off.get(rank.get(ext, ""), 0), ext, 0))
        name_data += name.get(ext, "").encode() + b"\0"
    out = bytearray(b"K2TAXDAT")
    out += struct.pack("<QQQ", len(nodes), len(name_data), len(rank_data))
    for n in nodes:
        out += struct.pack("<7Q", *n)
    out += name_data
    out += rank_data
    return bytes(out)


def write_k2_taxonomy(nodes_path: str, names_path: str, mapped_taxids, out_path: str) -> None:
    parent, rank, name = parse_dmp(nodes_path, names_path)
    with open(out_path, "wb") as f:
        f.write(build_k2_taxonomy(parent, rank, name, mapped_taxids))


def read_k2_taxonomy(path: str):
    """(parent_ids, external_ids) lists of a taxo.k2d."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != b"K2TAXDAT":
        raise ValueError(f"malformed taxonomy file {path}")
    n, _, _ = struct.unpack_from("<QQQ", b, 8)
    par, ext = [], []
    for i in range(n):
        rec = struct.unpack_from("<7Q", b, 32 + 56 * i)
        par.append(rec[0])
        ext.append(rec[5])
    return par, ext
