// Source of the batched shared-scan kernel.
//
// Shared scan: the fact columns touched by any query of the batch are
// streamed from HBM once per tile (128-bit nontemporal loads), then every
// template's pipeline body runs over those registers for each of its
// queries.  A template is a plan signature; its queries differ only in their
// parameter blocks (constants, hash tables, result slots), so the module
// depends on the set of templates, not on the batch size or constants.
#include <algorithm>
#include <map>
#include <sstream>

#include "hip_codegen.h"

namespace db {

namespace {

std::string withTemplate(const std::string& code, size_t t) {
    std::string out;
    const std::string tag = "@T@", idx = std::to_string(t);
    size_t at = 0;
    for (size_t k = code.find(tag); k != std::string::npos; k = code.find(tag, at)) {
        out += code.substr(at, k - at) + idx;
        at = k + tag.size();
    }
    return out + code.substr(at);
}

// pointers to query q's LDS aggregation region of template t (stride and
// hash size set by the launch)
std::string ldsRegion(const QueryProgram* p, size_t t, const std::string& ind) {
    std::ostringstream o;
    o << ind << "u64* s_agg = (u64*)((char*)cq_lds + B.lds_off[" << t << "] + (unsigned)(q - B.tq_begin[" << t
      << "]) * B.lds_stride[" << t << "]);\n"
      << ind << "u32* s_flag = (u32*)(s_agg + " << p->groups * (p->storage_ts - p->ngroups) << ");\n"
      << ind << "const unsigned cq_hbits = B.hbits[" << t << "];\n"
      << ind << "(void)s_flag; (void)cq_hbits;\n";
    return o.str();
}

}  // namespace

std::string generateBatchScanSource(const std::vector<const QueryProgram*>& templates, int bt, int vpt,
                                    std::vector<std::string>& union_cols, std::vector<bool>& union_nullable,
                                    unsigned* table_lds_bytes, unsigned table_mask, int* lds_candidates,
                                    unsigned table_lds_limit) {
    if (table_lds_limit > kBatchTableLdsMax) throw std::runtime_error("batch table LDS beyond its maximum");
    if (templates.empty()) throw std::runtime_error("batch without templates");
    if (templates.size() > 16) throw std::runtime_error("too many query templates in one batch (max 16)");
    union_cols.clear();
    union_nullable.clear();
    std::map<std::string, int> slot;
    for (const QueryProgram* t : templates) {
        if (t->batch_body.empty()) throw std::runtime_error("template generated without batch body");
        for (size_t s = 0; s < t->fact_cols.size(); ++s) {
            if (slot.count(t->fact_cols[s])) continue;
            slot[t->fact_cols[s]] = static_cast<int>(union_cols.size());
            union_cols.push_back(t->fact_cols[s]);
            union_nullable.push_back(t->fact_nullable[s]);
        }
    }
    if (union_cols.size() > 24) throw std::runtime_error("batch touches more than 24 fact columns");
    const int ipt = 4 * vpt;
    std::ostringstream src;
    src << "// batched shared scan over " << templates.size() << " templates\n"
        << "#include \"crystal_hip.h\"\n\n"
        << "#define BT " << bt << "\n#define VPT " << vpt << "\n#define IPT " << ipt << "\n#define TILE (BT * IPT)\n\n"
        << "extern \"C\" __global__ __launch_bounds__(BT) void cq_batch_scan(const cq_batch_params B) {\n"
        << "    using namespace chip;\n"
        << "    const int tid = threadIdx.x;\n";
    // per-query LDS aggregation regions (dynamic shared memory)
    bool any_lds = false;
    for (const QueryProgram* t : templates) any_lds = any_lds || t->batch_lds_stride > 0;
    if (any_lds) {
        src << "    extern __shared__ u64 cq_lds[];\n";
        for (size_t t = 0; t < templates.size(); ++t) {
            const QueryProgram* p = templates[t];
            if (p->batch_lds_stride == 0) continue;
            src << "    for (int q = B.tq_begin[" << t << "]; q < B.tq_end[" << t << "]; ++q) {\n"
                << ldsRegion(p, t, "        ")
                << p->batch_lds_init << "    }\n";
        }
        src << "    __syncthreads();\n";
    }
    // Small shared dimension tables live in LDS for the whole launch (one copy
    // per layout key: templates with equal keys share it), staged from the
    // first template with queries that owns one; larger ones stay in HBM.
    struct LdsTable {
        std::string key;
        uint64_t words = 0;
        std::vector<std::pair<size_t, size_t>> users;  // (template, join)
    };
    // query-mask tables of mask-form templates first (they replace the
    // per-query loop), within the same static LDS budget
    uint64_t mask_bytes = 0;
    std::vector<bool> use_mask(templates.size(), false);
    for (size_t t = 0; t < templates.size(); ++t) {
        const QueryProgram* p = templates[t];
        if (!p->batch_mask) continue;
        uint64_t b = 0;
        for (const auto& tb : p->batch_mask_tables) b += 8 * static_cast<uint64_t>(tb.second);
        if (mask_bytes + b > table_lds_limit) continue;
        mask_bytes += b;
        use_mask[t] = true;
        for (const auto& tb : p->batch_mask_tables)
            src << "    __shared__ u64 " << withTemplate(tb.first, t) << "[" << tb.second << "];\n";
    }
    for (size_t t = 0; t < templates.size(); ++t)
        if (use_mask[t]) src << withTemplate(templates[t]->batch_mask_build, t);
    // a launch of one mask-form template takes its deferred-probe body and
    // makes those probes inside it
    auto useDeferred = [&](size_t t) { return use_mask[t] && templates.size() == 1 && !templates[t]->batch_mask_body_deferred.empty(); };
    auto deferred = [&](size_t t, size_t j) {
        const auto& d = templates[t]->batch_mask_deferred;
        return useDeferred(t) && std::find(d.begin(), d.end(), static_cast<int>(j)) != d.end();
    };
    std::vector<LdsTable> cands;
    for (size_t t = 0; t < templates.size(); ++t) {
        const QueryProgram* p = templates[t];
        if (!p->shared_tables) continue;
        for (size_t j = 0; j < p->joins.size(); ++j) {
            const JoinProgram& jp = p->joins[j];
            if (!jp.shared || jp.db == 0 || jp.s_key.empty() || 4 * jp.ht_words > table_lds_limit) continue;
            size_t k = 0;
            while (k < cands.size() && cands[k].key != jp.s_key) ++k;
            if (k == cands.size()) cands.push_back(LdsTable{jp.s_key, jp.ht_words, {}});
            cands[k].users.emplace_back(t, j);
        }
    }
    // most templates served first, then the smallest
    std::stable_sort(cands.begin(), cands.end(), [](const LdsTable& x, const LdsTable& y) {
        return x.users.size() != y.users.size() ? x.users.size() > y.users.size() : x.words < y.words;
    });
    if (lds_candidates) *lds_candidates = static_cast<int>(cands.size());
    std::vector<LdsTable> tabs;
    uint64_t tab_bytes = 0;
    std::map<std::pair<size_t, size_t>, size_t> tab_of;
    for (size_t k = 0; k < cands.size(); ++k) {
        auto& c = cands[k];
        if (k < 32 && !((table_mask >> k) & 1u)) continue;  // placed in HBM
        if (mask_bytes + tab_bytes + 4 * c.words > table_lds_limit) continue;
        tab_bytes += 4 * c.words;
        for (const auto& u : c.users) tab_of[u] = tabs.size();
        tabs.push_back(std::move(c));
    }
    for (size_t k = 0; k < tabs.size(); ++k) src << "    __shared__ u32 cq_tab" << k << "[" << tabs[k].words << "];\n";
    for (size_t k = 0; k < tabs.size(); ++k) {
        src << "    {\n        const u32* src = nullptr;\n";
        for (size_t u = 0; u < tabs[k].users.size(); ++u) {
            const size_t t = tabs[k].users[u].first, j = tabs[k].users[u].second;
            src << "        " << (u ? "else " : "") << "if (B.tq_begin[" << t << "] < B.tq_end[" << t << "]) src = B.qp[B.tq_begin[" << t
                << "]].ht[" << j << "];\n";
        }
        src << "        if (src)\n            for (int w = tid; w < " << tabs[k].words << "; w += BT) cq_tab" << k << "[w] = src[w];\n"
            << "    }\n";
    }
    if (!tabs.empty() || mask_bytes > 0) src << "    __syncthreads();\n";
    if (table_lds_bytes) *table_lds_bytes = static_cast<unsigned>(tab_bytes + mask_bytes);

    // Launch-level probes: a (layout key, fact key column) that several
    // templates of the launch probe alike (equal keys: equal table
    // contents) is probed once per tile; those templates read its entries.
    // One random L2 gather per row instead of one per template.
    struct LaunchProbe {
        std::string key, fk, dim, text;
        bool fk_null = false;
        bool zmarks = false;  // the text marks probe key 0
        std::vector<std::pair<size_t, size_t>> users;  // (template, join)
    };
    std::vector<LaunchProbe> lps;
    for (size_t t = 0; t < templates.size(); ++t) {
        const QueryProgram* p = templates[t];
        if (!p->shared_tables || p->batch_shared_parts.size() != p->joins.size()) continue;
        for (size_t j = 0; j < p->joins.size(); ++j) {
            const JoinProgram& jp = p->joins[j];
            if (!jp.shared || jp.s_probe.empty() || deferred(t, j)) continue;
            // Probe key 0 (the reference's empty-slot corner): a template
            // whose every dimension key is in the table resolves it by the
            // plain lookup (zmode 0), others mark it (z) and read the entry
            // of the key sharing its slot.  Where that key lies outside the
            // key range (zmode 1, 2) one probe text serves both kinds: a
            // zmode-0 template reads key 0 as absent (m false) either way.
            // Where key 0 is itself a key (zmode 3) the marked form hides the
            // plain match, so those templates probe apart from the others.
            const std::string key = jp.s_key + "#" + jp.fk_col + (jp.zmode == 3 ? "#z3" : "");
            size_t k = 0;
            while (k < lps.size() && lps[k].key != key) ++k;
            if (k == lps.size()) {
                LaunchProbe lp;
                lp.key = key;
                lp.fk = jp.fk_col;
                lp.dim = jp.dim_table;
                const auto it = std::find(p->fact_cols.begin(), p->fact_cols.end(), jp.fk_col);
                lp.fk_null = it != p->fact_cols.end() && p->fact_nullable[static_cast<size_t>(it - p->fact_cols.begin())];
                lps.push_back(std::move(lp));
            }
            // the text of a template that marks key 0, when one does
            if (lps[k].text.empty() || (!lps[k].zmarks && jp.zmode != 0)) {
                lps[k].text = jp.s_probe;
                lps[k].zmarks = jp.zmode != 0;
            }
            lps[k].users.emplace_back(t, j);
        }
    }
    lps.erase(std::remove_if(lps.begin(), lps.end(), [](const LaunchProbe& lp) { return lp.users.size() < 2; }), lps.end());
    std::map<std::pair<size_t, size_t>, size_t> lp_of;
    for (size_t g = 0; g < lps.size(); ++g)
        for (const auto& u : lps[g].users) lp_of[u] = g;
    auto subst = [](std::string text, const std::string& from, const std::string& to) {
        for (size_t at = text.find(from); at != std::string::npos; at = text.find(from, at + to.size())) text.replace(at, from.size(), to);
        return text;
    };
    auto anyQueries = [](const std::vector<std::pair<size_t, size_t>>& users) {
        std::string c;
        for (size_t u = 0; u < users.size(); ++u)
            c += (u ? " || " : "") + std::string("B.tq_begin[") + std::to_string(users[u].first) + "] < B.tq_end[" +
                 std::to_string(users[u].first) + "]";
        return c;
    };

    // the last template (kernel order) that reads each union column: its
    // load is nontemporal, earlier templates' loads keep the tile in cache
    std::map<std::string, size_t> last_use;
    for (size_t t = 0; t < templates.size(); ++t)
        for (const auto& c : templates[t]->fact_cols) last_use[c] = t;
    src << "    for (long long tile = blockIdx.x; tile < B.num_tiles; tile += gridDim.x) {\n"
        << "        const long long base = B.row_begin + tile * (long long)TILE;\n"
        << "        const bool full = base + TILE <= B.row_end;\n"
        << "        u32 tile_fl = " << (ipt == 32 ? "0xffffffffu" : "((1u << IPT) - 1u)") << ";\n"
        << "        if (!full) {\n"
        << "            tile_fl = 0u;\n"
        << "#pragma unroll\n            for (int i = 0; i < IPT; ++i)\n"
        << "                if (base + (long long)(i >> 2) * (BT * 4) + tid * 4 + (i & 3) < B.row_end) tile_fl |= 1u << i;\n"
        << "        }\n";
    for (size_t g = 0; g < lps.size(); ++g) {
        const LaunchProbe& lp = lps[g];
        const std::string G = std::to_string(g);
        const int u = slot[lp.fk];
        src << "        // launch-level probe " << lp.dim << " on " << lp.fk << " (templates";
        for (const auto& us : lp.users) src << " " << us.first;
        src << ")\n"
            << "        u32 ge" << G << "[IPT]; bool gm" << G << "[IPT], gz" << G << "[IPT];\n"
            << "        if (" << anyQueries(lp.users) << ") {\n"
            << "        int gk" << G << "[IPT];\n";
        if (lp.fk_null) src << "        u32 gv" << G << " = 0u;\n";
        src << "        if (full) {\n"
            << "#pragma unroll\n            for (int v = 0; v < VPT; ++v) {\n"
            << "                const long long r = base + (long long)v * (BT * 4) + tid * 4;\n"
            << "                { const v4i x = *((const v4i*)(B.col[" << u << "] + r)); gk" << G << "[4 * v] = x.x; gk" << G
            << "[4 * v + 1] = x.y; gk" << G << "[4 * v + 2] = x.z; gk" << G << "[4 * v + 3] = x.w; }\n";
        if (lp.fk_null) src << "                gv" << G << " |= ((B.valid[" << u << "][r >> 5] >> (r & 31)) & 0xfu) << (4 * v);\n";
        src << "            }\n"
            << "        } else {\n"
            << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) {\n"
            << "                const long long r = base + (long long)(i >> 2) * (BT * 4) + tid * 4 + (i & 3);\n"
            << "                const bool in = r < B.row_end;\n"
            << "                gk" << G << "[i] = in ? B.col[" << u << "][r] : 0;\n";
        if (lp.fk_null) src << "                if (in) gv" << G << " |= chip::bit32(B.valid[" << u << "], r) << i;\n";
        src << "            }\n"
            << "        }\n"
            << "        const u32* gt" << G << " = nullptr;\n";
        for (size_t k = 0; k < lp.users.size(); ++k) {
            const size_t t = lp.users[k].first, j = lp.users[k].second;
            const auto it = tab_of.find({t, j});
            src << "        " << (k ? "else " : "") << "if (B.tq_begin[" << t << "] < B.tq_end[" << t << "]) gt" << G << " = "
                << (it != tab_of.end() ? "cq_tab" + std::to_string(it->second) : "B.qp[B.tq_begin[" + std::to_string(t) + "]].ht[" + std::to_string(j) + "]")
                << ";\n";
        }
        std::string text = lp.text;
        text = subst(text, "@FVTEST@", lp.fk_null ? "            if (!((gv" + G + " >> i) & 1u)) continue;\n" : "");
        text = subst(text, "@X@", "gx" + G);
        text = subst(text, "@M@", "gm" + G);
        text = subst(text, "@Z@", "gz" + G);
        text = subst(text, "@E@", "ge" + G);
        text = subst(text, "@FK@", "gk" + G + "[i]");
        text = subst(text, "@TAB@", "gt" + G);
        src << text << "        }\n";
    }
    for (size_t t = 0; t < templates.size(); ++t) {
        const QueryProgram* p = templates[t];
        src << "        // template " << t << ": " << p->signature << "\n"
            << "        if (B.tq_begin[" << t << "] < B.tq_end[" << t << "]) {\n";
        // this template's columns, loaded in its own scope: registers of
        // different templates do not overlap
        const size_t nf = p->fact_cols.size();
        for (size_t s2 = 0; s2 < nf; ++s2) {
            src << "        int c" << s2 << "[IPT];\n";
            if (p->fact_nullable[s2]) src << "        u32 v" << s2 << " = 0u;\n";
        }
        src << "        if (full) {\n"
            << "#pragma unroll\n            for (int v = 0; v < VPT; ++v) {\n"
            << "                const long long r = base + (long long)v * (BT * 4) + tid * 4;\n";
        for (size_t s2 = 0; s2 < nf; ++s2) {
            const int u = slot[p->fact_cols[s2]];
            const bool nt = last_use[p->fact_cols[s2]] == t;
            src << "                { const v4i x = " << (nt ? "__builtin_nontemporal_load(" : "*(") << "(const v4i*)(B.col[" << u
                << "] + r)); c" << s2 << "[4 * v] = x.x; c" << s2 << "[4 * v + 1] = x.y; c" << s2 << "[4 * v + 2] = x.z; c" << s2
                << "[4 * v + 3] = x.w; }\n";
            if (p->fact_nullable[s2])
                src << "                v" << s2 << " |= ((B.valid[" << u << "][r >> 5] >> (r & 31)) & 0xfu) << (4 * v);\n";
        }
        src << "            }\n"
            << "        } else {\n"
            << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) {\n"
            << "                const long long r = base + (long long)(i >> 2) * (BT * 4) + tid * 4 + (i & 3);\n"
            << "                const bool in = r < B.row_end;\n";
        for (size_t s2 = 0; s2 < nf; ++s2) {
            const int u = slot[p->fact_cols[s2]];
            src << "                c" << s2 << "[i] = in ? B.col[" << u << "][r] : 0;\n";
            if (p->fact_nullable[s2]) src << "                if (in) v" << s2 << " |= chip::bit32(B.valid[" << u << "], r) << i;\n";
        }
        src << "            }\n"
            << "        }\n";
        if (!p->batch_shared.empty()) {
            src << "        const cq_query_params& P0 = B.qp[B.tq_begin[" << t << "]];\n";
            for (size_t j = 0; j < p->joins.size(); ++j) {
                const auto it = tab_of.find({t, j});
                src << "        const u32* cq_sht" << j << " = " << (it != tab_of.end() ? "cq_tab" + std::to_string(it->second) : "P0.ht[" + std::to_string(j) + "]")
                    << ";\n";
            }
            for (size_t j = 0; j < p->joins.size(); ++j) {
                if (deferred(t, j)) continue;  // the mask body probes it
                const auto it = lp_of.find({t, j});
                if (it == lp_of.end()) {
                    src << (p->batch_shared_parts.size() == p->joins.size() ? p->batch_shared_parts[j] : (j ? std::string() : p->batch_shared));
                    continue;
                }
                const std::string J = std::to_string(j), G = std::to_string(it->second);
                src << "        // " << p->joins[j].dim_table << " on " << p->joins[j].fk_col << ": the launch-level probe's entries\n"
                    << "        u32 (&e" << J << ")[IPT] = ge" << G << "; bool (&m" << J << ")[IPT] = gm" << G << "; bool (&z" << J
                    << ")[IPT] = gz" << G << ";\n";
            }
        }
        if (use_mask[t]) {
            src << withTemplate(useDeferred(t) ? p->batch_mask_body_deferred : p->batch_mask_body, t) << "        }\n";
            continue;
        }
        src << "        for (int q = B.tq_begin[" << t << "]; q < B.tq_end[" << t << "]; ++q) {\n"
            << "            const cq_query_params& P = B.qp[q];\n";
        if (p->batch_lds_stride > 0) src << ldsRegion(p, t, "            ");
        if (p->batch_bool_flags)
            src << "            bool fl[IPT];\n#pragma unroll\n            for (int i = 0; i < IPT; ++i) fl[i] = (tile_fl >> i) & 1u;\n";
        else
            src << "            u32 fl = tile_fl;\n";
        src << "            do {\n"
            << p->batch_body
            << "            } while (0);\n"
            << "        }\n"
            << "        }\n";
    }
    src << "    }\n";
    if (any_lds) {
        // merge every query's LDS partials into its global slots
        src << "    __syncthreads();\n";
        for (size_t t = 0; t < templates.size(); ++t) {
            const QueryProgram* p = templates[t];
            if (p->batch_lds_stride == 0) continue;
            src << "    for (int q = B.tq_begin[" << t << "]; q < B.tq_end[" << t << "]; ++q) {\n"
                << "        const cq_query_params& P = B.qp[q];\n"
                << ldsRegion(p, t, "        ")
                << p->batch_lds_flush << "    }\n";
        }
    }
    src << "}\n";
    return src.str();
}

}  // namespace db
