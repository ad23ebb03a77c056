// NCBI nodes.dmp / names.dmp -> taxo.k2d (src/taxonomy.cc:19-185, 275-287):
// the taxonomy restricted to the root and the ancestors of mapped taxa, ids
// assigned breadth first from the root with each node's children taken in
// ascending external-id order (so a parent's id is below its children's and
// siblings are contiguous), then written in the Kraken 2 layout.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <set>

#include "k2g_host.h"

namespace k2host {

namespace {

// Splits a .dmp line on "\t|\t" after dropping the trailing "\t|".
std::vector<std::string> dmp_fields(std::string line, size_t max_fields) {
  if (!line.empty() && line.back() == '\r') line.pop_back();
  if (line.size() >= 2 && line.compare(line.size() - 2, 2, "\t|") == 0) line.resize(line.size() - 2);
  std::vector<std::string> out;
  size_t pos = 0;
  while (out.size() + 1 < max_fields) {
    const size_t q = line.find("\t|\t", pos);
    if (q == std::string::npos) break;
    out.push_back(line.substr(pos, q - pos));
    pos = q + 3;
  }
  out.push_back(line.substr(pos));
  return out;
}

}  // namespace

bool NcbiTaxonomy::load(const std::string &nodes_path, const std::string &names_path,
                        std::string &err) {
  std::ifstream nodes(nodes_path);
  if (!nodes) {
    err = "error opening " + nodes_path;
    return false;
  }
  std::string line;
  while (std::getline(nodes, line)) {
    auto f = dmp_fields(line, 4);
    if (f.size() < 3) continue;
    const uint64_t node = std::strtoull(f[0].c_str(), nullptr, 10);
    const uint64_t par = std::strtoull(f[1].c_str(), nullptr, 10);
    if (node == 0) {
      err = "attempt to create taxonomy w/ node ID == 0";
      return false;
    }
    parent[node] = node == 1 ? 0 : par;
    rank[node] = f[2];
    children[node == 1 ? 0 : par].push_back(node);
  }
  std::ifstream names(names_path);
  if (!names) {
    err = "error opening " + names_path;
    return false;
  }
  while (std::getline(names, line)) {
    auto f = dmp_fields(line, 5);
    if (f.size() < 4) continue;
    if (f[3] == "scientific name") name[std::strtoull(f[0].c_str(), nullptr, 10)] = f[1];
  }
  marked.insert(1);
  return true;
}

void NcbiTaxonomy::mark(uint64_t taxid) {
  while (taxid && !marked.count(taxid)) {
    marked.insert(taxid);
    auto it = parent.find(taxid);
    taxid = it == parent.end() ? 0 : it->second;
  }
}

bool NcbiTaxonomy::write_kraken(const std::string &path, std::string &err) const {
  std::set<std::string> known_ranks;
  for (auto &kv : rank) known_ranks.insert(kv.second);
  std::map<std::string, uint64_t> rank_off;
  std::string rank_data;
  for (auto &r : known_ranks) {
    rank_off[r] = rank_data.size();
    rank_data += r;
    rank_data.push_back('\0');
  }
  std::vector<Taxonomy::Node> nodes(1, Taxonomy::Node{0, 0, 0, 0, 0, 0, 0});
  std::string name_data;
  std::map<uint64_t, uint64_t> ext_to_int{{0, 0}};
  std::deque<uint64_t> queue{1};
  uint64_t next_id = 0;
  while (!queue.empty()) {
    const uint64_t ext = queue.front();
    queue.pop_front();
    const uint64_t id = ++next_id;
    ext_to_int[ext] = id;
    Taxonomy::Node n{};
    auto pit = parent.find(ext);
    const uint64_t ext_parent = pit == parent.end() ? 0 : pit->second;
    n.parent_id = ext_to_int.count(ext_parent) ? ext_to_int[ext_parent] : 0;
    n.first_child = id + 1 + queue.size();
    auto cit = children.find(ext);
    if (cit != children.end()) {
      std::vector<uint64_t> kids = cit->second;
      std::sort(kids.begin(), kids.end());
      for (uint64_t c : kids)
        if (marked.count(c)) {
          queue.push_back(c);
          n.child_count++;
        }
    }
    n.name_offset = name_data.size();
    auto rit = rank.find(ext);
    n.rank_offset = rit == rank.end() ? 0 : rank_off[rit->second];
    n.external_id = ext;
    nodes.push_back(n);
    auto nit = name.find(ext);
    if (nit != name.end()) name_data += nit->second;
    name_data.push_back('\0');
  }
  std::FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "unable to open " + path + " for writing";
    return false;
  }
  const uint64_t hdr[3] = {nodes.size(), name_data.size(), rank_data.size()};
  bool ok = std::fwrite("K2TAXDAT", 1, 8, f) == 8 && std::fwrite(hdr, 8, 3, f) == 3 &&
            std::fwrite(nodes.data(), sizeof(Taxonomy::Node), nodes.size(), f) == nodes.size() &&
            std::fwrite(name_data.data(), 1, name_data.size(), f) == name_data.size() &&
            std::fwrite(rank_data.data(), 1, rank_data.size(), f) == rank_data.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) err = "error writing " + path;
  return ok;
}

// src/build_db.cc:113-148: every sequence id of a header ('\x01' starts
// another id; whitespace ends the current one).
std::vector<std::string> sequence_ids(const std::string &header) {
  std::vector<std::string> list;
  std::string cur;
  bool in_id = true;
  for (char ch : header) {
    if (ch == 0x01) {
      if (!cur.empty()) list.push_back(cur);
      cur.clear();
      in_id = true;
    } else if (in_id && std::isspace((unsigned char) ch)) {
      if (!cur.empty()) list.push_back(cur);
      cur.clear();
      in_id = false;
    } else if (in_id) {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) list.push_back(cur);
  return list;
}

// src/utilities.cc:14-27: each template bit becomes bits_per_char bits.
uint64_t expand_spaced_seed_mask(uint64_t mask, int bits_per_char) {
  uint64_t out = 0;
  const uint64_t fill = (1ULL << bits_per_char) - 1;
  for (int i = 64 / bits_per_char - 1; i >= 0; i--) {
    out <<= bits_per_char;
    if ((mask >> i) & 1) out |= fill;
  }
  return out;
}

}  // namespace k2host
