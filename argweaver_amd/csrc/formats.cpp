// formats.cpp -- the reference's on-disk formats on the host: local trees
// (.smc: NAMES / REGION / TREE newick / SPR lines), site alignments (.sites:
// NAMES / REGION / one line per variant site) and the alignment compression
// that maps every c bases to one column while keeping every variant site.
//   .smc write  local_tree.cpp:1286-1375   .smc read  local_tree.cpp:1430-1565
//   .sites read sequences.cpp:173-313     to sequences sequences.cpp:315-343
//   compression sequences.cpp:523-626, sequences.h SitesMapping,
//               local_tree.cpp:790-873
// Files ending in .gz are read through zlib (the reference pipes them
// through gzip); plain files are read the same way.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "local_trees.h"

namespace aw {

struct Sites {
    std::vector<std::string> names;
    std::string chrom = "chr";
    int start_coord = 0, end_coord = 0;
    std::vector<int> positions;
    std::vector<std::string> cols;
    int length() const { return end_coord - start_coord; }
};

// SitesMapping (sequences.h): the old and new coordinates of the variant
// sites and the original position of every compressed column
struct SitesMapping {
    int old_start = 0, old_end = 0, new_start = 0, new_end = 0;
    std::vector<int> old_sites, new_sites, all_sites;

    void compress_blocks(const std::vector<int> &in, std::vector<int> &out) const
    {
        int cur = new_start;
        const int new_seqlen = new_end - new_start;
        int end = old_start;
        for (int len : in) {
            end += len;
            if (end < old_end) {
                int cur2 = cur;
                while (cur2 < new_seqlen && all_sites[cur2] < end) cur2++;
                out.push_back(cur2 - cur);
                cur = cur2;
            } else {
                out.push_back(new_end - cur);
            }
        }
    }
    void uncompress_blocks(const std::vector<int> &in, std::vector<int> &out) const
    {
        int cur = old_start, end = new_start;
        for (int len : in) {
            end += len;
            if (end < new_end) {
                // block ends placed at the midpoint between neighbouring columns
                const int cur2 = (all_sites[end - 1] + 1 + all_sites[end]) / 2;
                out.push_back(cur2 - cur);
                cur = cur2;
            } else {
                out.push_back(old_end - cur);
            }
        }
    }
};

namespace {

// line reader over a plain or gzip file
struct LineReader {
    gzFile f = nullptr;
    std::string line;
    explicit LineReader(const char *path) { f = gzopen(path, "rb"); }
    ~LineReader()
    {
        if (f) gzclose(f);
    }
    bool ok() const { return f != nullptr; }
    // next line without its end-of-line characters; false at the end
    bool next()
    {
        line.clear();
        char buf[65536];
        bool any = false;
        while (gzgets(f, buf, sizeof buf)) {
            any = true;
            line += buf;
            if (!line.empty() && line.back() == '\n') break;
        }
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        return any;
    }
};

std::vector<std::string> split_tabs(const char *s)
{
    std::vector<std::string> out;
    const char *p = s;
    while (true) {
        const char *q = strchr(p, '\t');
        if (!q) {
            out.emplace_back(p);
            break;
        }
        out.emplace_back(p, q - p);
        p = q + 1;
    }
    return out;
}

// closest time point (find_time, local_tree.cpp:1019-1034)
int nearest_time(double t, const double *times, int ntimes)
{
    double best = INFINITY;
    int bi = -1;
    for (int i = 0; i < ntimes; i++) {
        const double d = fabs(times[i] - t);
        if (d < best) {
            best = d;
            bi = i;
        }
    }
    return bi;
}

// "&&NHX:key=value:age=..." -> age (parse_node_age, local_tree.cpp:1067-1091)
bool nhx_age(const char *text, const char *end, double *age)
{
    if (end - text < 6 || strncmp(text, "&&NHX:", 6) != 0) return false;
    const char *key = text + 6;
    while (key < end) {
        const char *eq = key;
        while (eq < end && *eq != '=') eq++;
        if (eq >= end) return false;
        const char *val = eq + 1, *vend = val;
        while (vend < end && *vend != ':') vend++;
        if (eq - key == 3 && strncmp(key, "age", 3) == 0) {
            std::string v(val, vend - val);
            return sscanf(v.c_str(), "%lf", age) == 1;
        }
        key = vend + 1;
    }
    return false;
}

// newick with integer node names and NHX ages -> local tree
// (parse_local_tree, local_tree.cpp:1096-1230)
bool parse_newick(const char *nw, const double *times, int ntimes, LTree &t)
{
    const int len = (int)strlen(nw);
    std::vector<int> ptree{-1}, ages{-1}, names{-1}, stack;
    int node = 0;
    for (int i = 0; i < len; i++) {
        const char ch = nw[i];
        if (ch == '(') {
            ptree.push_back(node);
            ages.push_back(-1);
            names.push_back(-1);
            stack.push_back(node);
            node = (int)ptree.size() - 1;
        } else if (ch == ',') {
            if (stack.empty()) return false;
            ptree.push_back(stack.back());
            ages.push_back(-1);
            names.push_back(-1);
            node = (int)ptree.size() - 1;
        } else if (ch == ')') {
            if (stack.empty()) return false;
            node = stack.back();
            stack.pop_back();
        } else if (ch == ':' || ch == ';') {
        } else if (ch == '[') {
            int j = i + 1;
            while (j < len && nw[j] != ']') j++;
            if (j >= len) return false;
            double age;
            if (nhx_age(nw + i + 1, nw + j, &age)) ages[node] = nearest_time(age, times, ntimes);
            i = j;
        } else {
            const char last = i > 0 ? nw[i - 1] : '(';
            while (i < len && nw[i] == ' ') i++;
            int j = i;
            while (j < len && !strchr(")(,:;[", nw[j])) j++;
            if (last == ')' || last == '(' || last == ',') {
                if (sscanf(nw + i, "%d", &names[node]) != 1) return false;
            }
            i = j - 1;
        }
    }
    if (!stack.empty()) return false;
    const int n = (int)ptree.size();
    for (int x : names)
        if (x < 0 || x >= n) return false;
    t.nnodes = n;
    t.nodes.assign(n, LNode{-1, {-1, -1}, 0});
    t.root = -1;
    for (int i = 0; i < n; i++) {
        const int j = names[i];
        if (ptree[i] != -1) {
            t.nodes[j].parent = names[ptree[i]];
        } else {
            t.nodes[j].parent = -1;
            t.root = j;
        }
        t.nodes[j].age = ages[i];
    }
    for (int i = 0; i < n; i++) {
        if (ptree[i] == -1) continue;
        int *c = t.nodes[names[ptree[i]]].child;
        if (c[0] == -1) c[0] = names[i];
        else if (c[1] == -1) c[1] = names[i];
        else return false;   // not binary
    }
    for (int i = 0; i < n; i++)
        if (t.nodes[i].child[0] == -1 && t.nodes[i].age == -1) t.nodes[i].age = 0;
    return true;
}

void write_newick_node(FILE *out, const LTree &t, const std::vector<std::string> &ids,
                       const double *times, int node, int depth)
{
    const LNode &nd = t.nodes[node];
    const double dist = nd.parent == -1 ? 0.0 : times[t.nodes[nd.parent].age] - times[nd.age];
    if (nd.child[0] == -1) {
        fprintf(out, "%s:%f[&&NHX:age=%f]", ids[node].c_str(), dist, times[nd.age]);
        return;
    }
    fprintf(out, "(");
    write_newick_node(out, t, ids, times, nd.child[0], depth + 1);
    fprintf(out, ",");
    write_newick_node(out, t, ids, times, nd.child[1], depth + 1);
    fprintf(out, ")");
    if (depth > 0)
        fprintf(out, "%s:%f[&&NHX:age=%f]", ids[node].c_str(), dist, times[nd.age]);
    else
        fprintf(out, "%s[&&NHX:age=%f]", ids[node].c_str(), times[nd.age]);
}

}  // namespace

// read_local_trees (local_tree.cpp:1430-1545); nullptr on a malformed file
ArgTrees *read_smc(const char *path, const double *times, int ntimes,
                   std::vector<std::string> *names_out)
{
    LineReader in(path);
    if (!in.ok()) {
        fprintf(stderr, "argweaver_amd: cannot read file '%s'\n", path);
        return nullptr;
    }
    ArgTrees *tr = new ArgTrees();
    std::vector<std::string> names;
    Spr spr{-1, -1, -1, -1};
    int nnodes = 0, lineno = 1;
    const LTree *last = nullptr;
    auto fail = [&](const char *what) -> ArgTrees * {
        fprintf(stderr, "argweaver_amd: %s in '%s' (line %d)\n", what, path, lineno);
        delete tr;
        return nullptr;
    };
    while (in.next()) {
        const char *line = in.line.c_str();
        if (strncmp(line, "NAMES", 5) == 0) {
            names = split_tabs(line + (line[5] ? 6 : 5));
            nnodes = 2 * (int)names.size() - 1;
        } else if (strncmp(line, "RANGE", 5) == 0) {
            return fail("deprecated RANGE line (use REGION)");
        } else if (strncmp(line, "REGION\t", 7) == 0) {
            char chrom[51];
            if (sscanf(line + 7, "%50s\t%d\t%d", chrom, &tr->start_coord, &tr->end_coord) != 3)
                return fail("bad REGION line");
            tr->chrom = chrom;
            tr->start_coord--;
        } else if (strncmp(line, "TREE", 4) == 0) {
            int start, end;
            if (sscanf(line + 5, "%d\t%d", &start, &end) != 2) return fail("bad TREE line");
            const char *nw = strchr(line + 5, '\t');
            nw = nw ? strchr(nw + 1, '\t') : nullptr;
            if (!nw) return fail("bad TREE line");
            LBlock B;
            if (!parse_newick(nw + 1, times, ntimes, B.tree)) return fail("bad newick");
            if (nnodes == 0) nnodes = B.tree.nnodes;
            if (B.tree.nnodes != nnodes) return fail("tree size differs from NAMES");
            B.spr = spr;
            if (!spr_is_null(spr)) {
                B.mapping.resize(nnodes);
                for (int i = 0; i < nnodes; i++) B.mapping[i] = i;
                B.mapping[last->nodes[spr.rn].parent] = -1;
            }
            B.blocklen = end - start + 1;
            tr->blocks.push_back(std::move(B));
            last = &tr->blocks.back().tree;
        } else if (strncmp(line, "SPR", 3) == 0) {
            int pos;
            double rt, ct;
            if (sscanf(line + 4, "%d\t%d\t%lf\t%d\t%lf", &pos, &spr.rn, &rt, &spr.cn, &ct) != 5)
                return fail("bad SPR line");
            spr.rt = nearest_time(rt, times, ntimes);
            spr.ct = nearest_time(ct, times, ntimes);
        }
        lineno++;
    }
    if (!tr->blocks.empty()) {
        tr->nnodes = tr->blocks.front().tree.nnodes;
        tr->seqids.resize(tr->num_leaves());
        for (int i = 0; i < tr->num_leaves(); i++) tr->seqids[i] = i;
    }
    if (!check_trees(*tr)) return fail("local trees fail the structural checks");
    if (names_out) *names_out = names;
    return tr;
}

// write_local_trees (local_tree.cpp:1286-1341): node names follow each node
// back to its name in the first tree through the block mappings
bool write_smc(const char *path, const ArgTrees &tr, const char *const *names,
               const double *times)
{
    FILE *out = fopen(path, "w");
    if (!out) {
        fprintf(stderr, "argweaver_amd: cannot write file '%s'\n", path);
        return false;
    }
    const int nnodes = tr.nnodes;
    if (names) {
        fprintf(out, "NAMES");
        for (int i = 0; i < tr.num_leaves(); i++) fprintf(out, "\t%s", names[tr.seqids[i]]);
        fprintf(out, "\n");
    }
    fprintf(out, "REGION\t%s\t%d\t%d\n", tr.chrom.c_str(), tr.start_coord + 1, tr.end_coord);
    std::vector<int> total(nnodes), tmp(nnodes);
    for (int i = 0; i < nnodes; i++) total[i] = i;
    std::vector<std::string> ids(nnodes);
    int end = tr.start_coord;
    for (size_t bi = 0; bi < tr.blocks.size(); bi++) {
        const LBlock &B = tr.blocks[bi];
        const int start = end;
        end += B.blocklen;
        for (int i = 0; i < nnodes; i++) ids[i] = std::to_string(total[i]);
        fprintf(out, "TREE\t%d\t%d\t", start + 1, end);
        write_newick_node(out, B.tree, ids, times, B.tree.root, 0);
        fprintf(out, ";\n");
        if (bi + 1 < tr.blocks.size()) {
            const LBlock &N = tr.blocks[bi + 1];
            const Spr &s = N.spr;
            fprintf(out, "SPR\t%d\t%d\t%f\t%d\t%f\n", end, total[s.rn], times[s.rt], total[s.cn],
                    times[s.ct]);
            // get_recoal_node (local_tree.cpp:473-490): the node nothing maps to
            std::vector<char> mapped(nnodes, 0);
            for (int i = 0; i < nnodes; i++)
                if (N.mapping[i] != -1) mapped[N.mapping[i]] = 1;
            int recoal = -1;
            for (int i = 0; i < nnodes && recoal < 0; i++)
                if (!mapped[i]) recoal = i;
            tmp = total;
            for (int i = 0; i < nnodes; i++) {
                if (N.mapping[i] != -1) total[N.mapping[i]] = tmp[i];
                else if (recoal >= 0) total[recoal] = tmp[i];
            }
        }
    }
    fclose(out);
    return true;
}

// read_sites (sequences.cpp:173-296); nullptr on a malformed file
Sites *read_sites_file(const char *path, int sub_start, int sub_end)
{
    LineReader in(path);
    if (!in.ok()) {
        fprintf(stderr, "argweaver_amd: cannot read file '%s'\n", path);
        return nullptr;
    }
    Sites *s = new Sites();
    int lineno = 0;
    auto fail = [&](const char *what) -> Sites * {
        fprintf(stderr, "argweaver_amd: %s in '%s' (line %d)\n", what, path, lineno);
        delete s;
        return nullptr;
    };
    while (in.next()) {
        lineno++;
        const char *line = in.line.c_str();
        if (strncmp(line, "NAMES\t", 6) == 0) {
            s->names = split_tabs(line + 6);
            for (const std::string &n : s->names)
                if (n.empty()) return fail("zero-length sequence name");
        } else if (strncmp(line, "REGION\t", 7) == 0) {
            char chrom[51];
            if (sscanf(line, "REGION\t%50s\t%d\t%d", chrom, &s->start_coord, &s->end_coord) != 3)
                return fail("bad REGION line");
            s->chrom = chrom;
            s->start_coord--;
            if (sub_start != -1) s->start_coord = sub_start;
            if (sub_end != -1) s->end_coord = sub_end;
        } else if (strncmp(line, "RANGE\t", 6) == 0) {
            return fail("deprecated RANGE line (use REGION)");
        } else {
            int pos;
            if (sscanf(line, "%d\t", &pos) != 1) return fail("first column is not an integer");
            // the 1-based position is compared with the 0-based region, as
            // the reference does
            if (pos < s->start_coord || pos >= s->end_coord) continue;
            const char *tab = strchr(line, '\t');
            if (!tab) return fail("site line without bases");
            const char *col = tab + 1;
            if (strlen(col) != s->names.size()) return fail("number of bases differs from NAMES");
            for (const char *c = col; *c; c++)
                if (!strchr("ACGTNacgtn", *c)) return fail("invalid sequence character");
            pos--;
            if (!s->positions.empty() && s->positions.back() >= pos)
                return fail("sites must be sorted and unique");
            s->positions.push_back(pos);
            s->cols.emplace_back(col);
        }
    }
    return s;
}

// find_compress_cols (sequences.cpp:523-599); false when the alignment
// cannot be compressed by that factor
bool find_compress_cols(const Sites &s, int compress, SitesMapping &m)
{
    const int ncols = (int)s.positions.size();
    m = SitesMapping();
    m.old_start = s.start_coord;
    m.old_end = s.end_coord;
    if (compress == 1) {
        for (int i = s.start_coord; i < s.end_coord; i++) m.all_sites.push_back(i);
        for (int i = 0; i < ncols; i++) {
            m.old_sites.push_back(s.positions[i]);
            m.new_sites.push_back(s.positions[i] - s.start_coord);
        }
        m.new_start = 0;
        m.new_end = s.length();
        return true;
    }
    int blocki = 0, next_block = s.start_coord + compress;
    const int half = compress / 2;
    for (int i = 0; i < ncols; i++) {
        const int col = s.positions[i];
        while (col >= next_block) {
            m.all_sites.push_back(next_block - half);
            next_block += compress;
            blocki++;
        }
        m.old_sites.push_back(col);
        m.new_sites.push_back(blocki);
        m.all_sites.push_back(col);
        next_block += compress;
        blocki++;
        if (next_block - compress > s.end_coord && i != ncols - 1) return false;
    }
    while (s.end_coord >= next_block) {
        m.all_sites.push_back(next_block - half);
        next_block += compress;
        blocki++;
    }
    m.new_start = 0;
    const int new_end = s.length() / compress;
    m.new_end = ncols > 0 ? std::max(m.new_sites[ncols - 1] + 1, new_end) : new_end;
    return true;
}

}  // namespace aw

using namespace aw;

extern "C" {

LocalTrees *read_local_trees(const char *filename, const double *times, int ntimes)
{
    return reinterpret_cast<LocalTrees *>(read_smc(filename, times, ntimes, nullptr));
}

void write_local_trees(char *filename, LocalTrees *trees, char **names, double *times, int ntimes)
{
    (void)ntimes;
    write_smc(filename, *reinterpret_cast<ArgTrees *>(trees), names, times);
}

void *arghmm_read_sites(const char *filename, int subregion_start, int subregion_end)
{
    return read_sites_file(filename, subregion_start, subregion_end);
}

void arghmm_delete_sites(void *sites) { delete reinterpret_cast<Sites *>(sites); }

int awamd_sites_shape(void *sites, int *shape4)
{
    const Sites &s = *reinterpret_cast<Sites *>(sites);
    shape4[0] = (int)s.names.size();
    shape4[1] = (int)s.positions.size();
    shape4[2] = s.start_coord;
    shape4[3] = s.end_coord;
    return 0;
}

const char *awamd_sites_name(void *sites, int i)
{
    const Sites &s = *reinterpret_cast<Sites *>(sites);
    return (i >= 0 && i < (int)s.names.size()) ? s.names[i].c_str() : nullptr;
}

int awamd_sites_data(void *sites, int *positions, char *cols)
{
    const Sites &s = *reinterpret_cast<Sites *>(sites);
    const size_t n = s.names.size();
    for (size_t i = 0; i < s.positions.size(); i++) {
        if (positions) positions[i] = s.positions[i];
        if (cols) memcpy(cols + i * n, s.cols[i].data(), n);
    }
    return (int)s.positions.size();
}

int awamd_sites_sequences(void *sites, char default_char, char *seqs)
{
    // make_sequences_from_sites (sequences.cpp:315-343): row i holds
    // sequence i over the region, default_char off the variant sites
    const Sites &s = *reinterpret_cast<Sites *>(sites);
    const int L = s.length(), nseqs = (int)s.names.size(), nsites = (int)s.positions.size();
    for (int i = 0; i < nseqs; i++) {
        char *row = seqs + (size_t)i * L;
        int col = 0;
        for (int j = 0; j < L; j++) {
            if (col < nsites && s.start_coord + j == s.positions[col]) row[j] = s.cols[col++][i];
            else row[j] = default_char;
        }
    }
    return L;
}

void *awamd_sites_compress(void *sites, int compress)
{
    // find_compress_cols + compress_sites (sequences.cpp:523-611)
    Sites &s = *reinterpret_cast<Sites *>(sites);
    if (compress < 1) return nullptr;
    SitesMapping *m = new SitesMapping();
    if (!find_compress_cols(s, compress, *m)) {
        fprintf(stderr, "argweaver_amd: the sites cannot be compressed by %d\n", compress);
        delete m;
        return nullptr;
    }
    s.start_coord = m->new_start;
    s.end_coord = m->new_end;
    for (size_t i = 0; i < s.positions.size(); i++) s.positions[i] = m->new_sites[i];
    return m;
}

int awamd_sites_uncompress(void *sites, void *mapping)
{
    // uncompress_sites (sequences.cpp:614-623)
    Sites &s = *reinterpret_cast<Sites *>(sites);
    const SitesMapping &m = *reinterpret_cast<SitesMapping *>(mapping);
    s.start_coord = m.old_start;
    s.end_coord = m.old_end;
    for (size_t i = 0; i < s.positions.size(); i++) s.positions[i] = m.old_sites[i];
    return 0;
}

void awamd_delete_sites_mapping(void *mapping) { delete reinterpret_cast<SitesMapping *>(mapping); }

int awamd_compress_local_trees(LocalTrees *trees, void *mapping)
{
    // compress_local_trees (local_tree.cpp:817-848), exact (not fuzzy)
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    const SitesMapping &m = *reinterpret_cast<SitesMapping *>(mapping);
    std::vector<int> in, out;
    for (const LBlock &B : tr.blocks) in.push_back(B.blocklen);
    m.compress_blocks(in, out);
    for (int x : out)
        if (x <= 0) {
            fprintf(stderr, "argweaver_amd: a block compresses to no column\n");
            return 1;
        }
    for (size_t i = 0; i < out.size(); i++) tr.blocks[i].blocklen = out[i];
    tr.start_coord = m.new_start;
    tr.end_coord = m.new_end;
    return 0;
}

int awamd_uncompress_local_trees(LocalTrees *trees, void *mapping)
{
    // uncompress_local_trees (local_tree.cpp:790-814)
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    const SitesMapping &m = *reinterpret_cast<SitesMapping *>(mapping);
    std::vector<int> in, out;
    for (const LBlock &B : tr.blocks) in.push_back(B.blocklen);
    m.uncompress_blocks(in, out);
    for (int x : out)
        if (x <= 0) {
            fprintf(stderr, "argweaver_amd: a block uncompresses to no base\n");
            return 1;
        }
    for (size_t i = 0; i < out.size(); i++) tr.blocks[i].blocklen = out[i];
    tr.start_coord = m.old_start;
    tr.end_coord = m.old_end;
    return 0;
}

}  // extern "C"
