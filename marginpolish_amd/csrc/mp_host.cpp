// mp_host.cpp -- params JSON loader, band geometry, traceback windows and
// large-gap split points for the pair-HMM engine (host side, no HIP).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "mp_internal.hpp"
#include "mp_json.hpp"

namespace mp {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

}  // namespace mp

extern "C" const char *mp_last_error(void) { return mp::g_error.c_str(); }
extern "C" const char *mp_version(void) { return "mp_pairhmm 0.1.0 gfx950"; }

namespace mp {

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------

static double lg(double v) { return v > 0.0 ? std::log(v) : -INFINITY; }

static void default_transitions(double *t) {
    // stateMachine3_construct defaults (stateMachine.c:597-606)
    t[kMatchContinue] = -0.030064059121770816;
    t[kMatchFromGapX] = -1.272871422049609;
    t[kMatchFromGapY] = -1.272871422049609;
    t[kGapOpenX] = -4.21256642;
    t[kGapOpenY] = -4.21256642;
    t[kGapExtendX] = -0.3388262689231553;
    t[kGapExtendY] = -0.3388262689231553;
    t[kGapSwitchToX] = -4.910694825551255;
    t[kGapSwitchToY] = -4.910694825551255;
}

static void default_emissions(Machine &m) {
    // setNucleotideEmissionMatchProbsToDefaults / GapProbsToDefaults (stateMachine.c:396-422):
    // match on the diagonal, A<->G and C<->T transitions, the rest transversions.
    const double kM = -1.8917761142, kTv = -4.3459578861, kTs = -3.760242452;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            bool transition = (i ^ j) == 2;  // A(0)<->G(2), C(1)<->T(3)
            m.match[i * 4 + j] = i == j ? kM : (transition ? kTs : kTv);
        }
    for (int i = 0; i < 4; i++) m.gapX[i] = m.gapY[i] = -1.3862943611;
}

static void default_band_params(mp_band_params &p) {
    // pairwiseAlignmentBandingParameters_construct (pairwiseAligner.c:999-1011)
    p.threshold = 0.01;
    p.minDiagsBetweenTraceBack = 1000;
    p.traceBackDiagonals = 40;
    p.diagonalExpansion = 20;
    p.splitMatrixBiggerThanThis = (int64_t)3000 * 3000;
    p.dynamicAnchorExpansion = 0;
}

// nucleotideEmissions_reverseComplement (stateMachine.c:454-470): M'[i][j] = M[3-i][3-j], G'[i] = G[3-i].
static Machine reverse_complement(const Machine &f) {
    Machine r = f;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.match[i * 4 + j] = f.match[(3 - i) * 4 + (3 - j)];
    for (int i = 0; i < 4; i++) {
        r.gapX[i] = f.gapX[3 - i];
        r.gapY[i] = f.gapY[3 - i];
    }
    return r;
}

using mpjson::Value;

[[noreturn]] static void bad(const std::string &msg) { throw std::runtime_error(msg); }

static double num(const Value &v, const char *what) {
    if (v.kind == Value::Number) return v.num;
    bad(std::string("expected a number for ") + what);
}

static bool boolean(const Value &v, const char *what) {
    if (v.kind == Value::Bool) return v.b;
    if (v.kind == Value::Number) return v.num != 0.0;
    bad(std::string("expected a boolean for ") + what);
}

static void floats(const Value &v, double *out, size_t n, const char *what) {
    if (v.kind != Value::Array || v.arr.size() != n)
        bad(std::string("expected an array of ") + std::to_string(n) + " numbers for " + what);
    for (size_t i = 0; i < n; i++) out[i] = num(v.arr[i], what);
}

// The 3-state slots of a five-state machine hold its short-gap sub-machine
// (M, short X, short Y), for inspection through mp_model_tables.
static void short_gap_slots(Machine &m) {
    const double *l5 = m.trans5;
    m.trans[kMatchContinue] = l5[0];
    m.trans[kGapOpenX] = l5[1];
    m.trans[kGapOpenY] = l5[2];
    m.trans[kMatchFromGapX] = l5[5];
    m.trans[kGapExtendX] = l5[6];
    m.trans[kGapSwitchToY] = l5[7];
    m.trans[kMatchFromGapY] = l5[10];
    m.trans[kGapSwitchToX] = l5[11];
    m.trans[kGapExtendY] = l5[12];
}

// hmm_jsonParse + hmm_getStateMachine (stateMachine.c:206-266, 632-689).
static Machine parse_hmm(const Value &h) {
    if (h.kind != Value::Object || h.obj.size() < 2 || h.obj[0].first != "type" || h.obj[1].first != "emissionsType")
        bad("hmm json must start with 'type' then 'emissionsType'");
    int type = (int)num(h.obj[0].second, "type");
    int etype = (int)num(h.obj[1].second, "emissionsType");
    if (type != 2 && type != 3 && type != 5) bad("unrecognised state machine type " + std::to_string(type));
    if (etype != 0) bad("unrecognised emission type " + std::to_string(etype));
    const int nt = type == 5 ? 25 : 9;
    double t[25], e[24];
    bool gotT = false, gotE = false;
    for (size_t i = 2; i < h.obj.size(); i++) {
        const std::string &k = h.obj[i].first;
        if (k == "transitions") {
            floats(h.obj[i].second, t, (size_t)nt, "transitions");
            gotT = true;
        } else if (k == "emissions") {
            floats(h.obj[i].second, e, 24, "emissions");
            gotE = true;
        } else if (k == "likelihood") {
            num(h.obj[i].second, "likelihood");
        } else {
            bad("unrecognised key in hmm json: " + k);
        }
    }
    if (!gotE) bad("did not find emissions in hmm json");
    if (!gotT) bad("did not find transitions in hmm json");
    Machine m;
    if (type == 5) {
        // Five-state extension (no counterpart in the reference, which knows
        // types 2 and 3 only): 25 probabilities [from * 5 + to] over
        // (M, short X, short Y, long X, long Y).
        for (int i = 0; i < 25; i++) {
            if (!(t[i] >= 0.0)) bad("five-state transitions must be non-negative probabilities");
            m.trans5[i] = lg(t[i]);
        }
        m.nStates = 5;
        short_gap_slots(m);
    }
    auto T = [&](int a, int b) { return t[a * 3 + b]; };
    const int M = 0, X = 1, Y = 2;
    if (type == 2) {  // symmetric: average the X and Y variants in probability space
        m.trans[kMatchContinue] = lg(T(M, M));
        m.trans[kMatchFromGapX] = m.trans[kMatchFromGapY] = lg((T(X, M) + T(Y, M)) / 2.0);
        m.trans[kGapOpenX] = m.trans[kGapOpenY] = lg((T(M, X) + T(M, Y)) / 2.0);
        m.trans[kGapExtendX] = m.trans[kGapExtendY] = lg((T(X, X) + T(Y, Y)) / 2.0);
        m.trans[kGapSwitchToX] = m.trans[kGapSwitchToY] = lg((T(Y, X) + T(X, Y)) / 2.0);
    } else if (type == 3) {
        m.trans[kMatchContinue] = lg(T(M, M));
        m.trans[kMatchFromGapX] = lg(T(X, M));
        m.trans[kMatchFromGapY] = lg(T(Y, M));
        m.trans[kGapOpenX] = lg(T(M, X));
        m.trans[kGapOpenY] = lg(T(M, Y));
        m.trans[kGapExtendX] = lg(T(X, X));
        m.trans[kGapExtendY] = lg(T(Y, Y));
        m.trans[kGapSwitchToX] = lg(T(Y, X));
        m.trans[kGapSwitchToY] = lg(T(X, Y));
    }
    for (int i = 0; i < 16; i++) m.match[i] = lg(e[i]);
    for (int i = 0; i < 4; i++) {
        m.gapX[i] = lg(e[16 + i]);
        m.gapY[i] = lg(e[20 + i]);
    }
    return m;
}

// repeatSubMatrix_jsonParse (parser.c:105-152): forward-strand keys only.
static void parse_repeat(const Value &v, std::vector<double> &tab) {
    tab.assign(kRepeatEntries, 0.0);
    if (v.kind != Value::Object) bad("repeatCountSubstitutionMatrix must be an object");
    std::vector<double> tmp(kMaxRepeat);
    for (auto &kv : v.obj) {
        const std::string &k = kv.first;
        if (k == "baseLogRepeatCounts_AT" || k == "baseLogRepeatCounts_GC") {
            floats(kv.second, tmp.data(), kMaxRepeat, k.c_str());
            continue;
        }
        if (k.size() != 31) bad("unrecognised key in repeat sub matrix json: " + k);
        const char *bases = "ACGT";
        const char *pb = std::strchr(bases, k[28]);
        if (!pb || k[28] == 0) bad("unrecognised base in repeat sub matrix json: " + k);
        if (k[30] != 'F') bad("unrecognised strand in repeat sub matrix json: " + k);
        int b = (int)(pb - bases);
        floats(kv.second, &tab[(size_t)b * kMaxRepeat * kMaxRepeat], (size_t)kMaxRepeat * kMaxRepeat, k.c_str());
    }
}

// pairwiseAlignmentParameters_jsonParse (pairwiseAligner.c:1017-1066).
static void parse_pairwise(const Value &v, mp_band_params &p) {
    default_band_params(p);
    if (v.kind != Value::Object) bad("pairwiseAlignmentParameters must be an object");
    for (auto &kv : v.obj) {
        const std::string &k = kv.first;
        if (k == "threshold") p.threshold = num(kv.second, "threshold");
        else if (k == "minDiagsBetweenTraceBack") p.minDiagsBetweenTraceBack = (int64_t)num(kv.second, "minDiags");
        else if (k == "traceBackDiagonals") p.traceBackDiagonals = (int64_t)num(kv.second, "traceBackDiagonals");
        else if (k == "diagonalExpansion") p.diagonalExpansion = (int64_t)num(kv.second, "diagonalExpansion");
        else if (k == "constraintDiagonalTrim") num(kv.second, "constraintDiagonalTrim");
        else if (k == "splitMatrixBiggerThanThis") p.splitMatrixBiggerThanThis = (int64_t)num(kv.second, "split");
        else if (k == "alignAmbiguityCharacters") boolean(kv.second, "alignAmbiguityCharacters");
        else if (k == "gapGamma") num(kv.second, "gapGamma");
        else if (k == "dynamicAnchorExpansion") p.dynamicAnchorExpansion = boolean(kv.second, "dynamic") ? 1 : 0;
        else bad("unrecognised key in pairwise alignment parameters json: " + k);
    }
    if (p.diagonalExpansion % 2 != 0) bad("pairwiseAlignmentParameters.diagonalExpansion must be even");
}

// polishParams_jsonParse (parser.c:154-408): the keys this path reads are
// interpreted; every other documented key is accepted and ignored; anything
// else is an error.
static void parse_polish(const Value &v, mp_model &m) {
    static const char *kKnown[] = {
        "useRunLengthEncoding", "referenceBasePenalty", "minPosteriorProbForAlignmentAnchors",
        "repeatCountSubstitutionMatrix", "poaConstructCompareRepeatCounts", "hmmForwardStrandReadGivenReference",
        "pairwiseAlignmentParameters", "shuffleChunks", "includeSoftClipping", "useRepeatCountsInAlignment",
        "chunkSize", "chunkBoundary", "maxDepth", "includeSecondaryAlignments", "includeSupplementaryAlignments",
        "filterAlignmentsWithMapQBelowThisThreshold", "candidateVariantWeight", "columnAnchorTrim",
        "maxConsensusStrings", "maxPoaConsensusIterations", "minPoaConsensusIterations",
        "maxRealignmentPolishIterations", "minRealignmentPolishIterations", "minReadsToCallConsensus",
        "filterReadsWhileHaveAtLeastThisCoverage", "minAvgBaseQuality", "hetSubstitutionProbability",
        "hetRunLengthSubstitutionProbability", "useReadAlleles", "useReadAllelesInPhasing", "alphabet"};
    if (v.kind != Value::Object) bad("polish params must be an object");
    bool useRLE = true, useRepeatCounts = false, gotHmm = false, gotPW = false, gotRep = false, gotAlpha = false;
    Machine fwd;
    for (auto &kv : v.obj) {
        const std::string &k = kv.first;
        bool known = false;
        for (const char *kk : kKnown)
            if (k == kk) known = true;
        if (!known) bad("unrecognised key in polish params json: " + k);
        if (k == "useRunLengthEncoding") useRLE = boolean(kv.second, k.c_str());
        else if (k == "useRepeatCountsInAlignment") useRepeatCounts = boolean(kv.second, k.c_str());
        else if (k == "hmmForwardStrandReadGivenReference") { fwd = parse_hmm(kv.second); gotHmm = true; }
        else if (k == "pairwiseAlignmentParameters") { parse_pairwise(kv.second, m.p); gotPW = true; }
        else if (k == "repeatCountSubstitutionMatrix") { parse_repeat(kv.second, m.repeatLog10); gotRep = true; }
        else if (k == "alphabet") {
            if (kv.second.kind != Value::String || kv.second.str != "nucleotide")
                bad("unrecognised alphabet type json");
            gotAlpha = true;
        }
    }
    if (!gotHmm) bad("did not find HMM for alignment of read to a reference in json polish params");
    if (!gotPW) bad("did not find pairwise alignment params in json polish params");
    if (!gotAlpha) bad("did not find alphabet params in json polish params");
    m.mach[1] = fwd;
    m.mach[0] = reverse_complement(fwd);
    m.mach[1].strand = 1;
    m.mach[0].strand = 0;
    if (!gotRep) m.repeatLog10.assign(kRepeatEntries, 0.0);
    if (useRepeatCounts) {
        if (!useRLE) bad("repeat counts in alignment requested without run length encoding");
        if (!gotRep) bad("repeat counts in alignment requested but no repeatCountSubstitutionMatrix");
        m.mach[0].rle = m.mach[1].rle = 1;  // rleNucleotideEmissions_construct (parser.c:391-401)
    }
}

// ---------------------------------------------------------------------------
// Band geometry
// ---------------------------------------------------------------------------

bool anchors_valid(const int64_t *a, int64_t n, int64_t lX, int64_t lY) {
    int64_t px = -1, py = -1;
    for (int64_t i = 0; i < n; i++) {
        int64_t x = a[3 * i], y = a[3 * i + 1], e = a[3 * i + 2];
        if (x <= px || y <= py || x >= lX || y >= lY || e < 0 || (e & 1)) return false;
        px = x;
        py = y;
    }
    return true;
}

static inline int64_t odd_up(int64_t xay, int64_t v) { return ((xay + v) & 1) ? v + 1 : v; }
static inline int64_t clampz(int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// The band of anti-diagonal xay inside box [xL,xU] x [yU,yL] (x >= xL,
// y <= yL, x <= xU, y >= yU), in closed form: the reference's four
// successive clamps (pairwiseAligner.c:97-113) reduce to a max and a min.
static inline bool box_diag(int64_t xay, int64_t xL, int64_t yL, int64_t xU, int64_t yU, int32_t &Lo, int32_t &Hi) {
    int64_t l = std::max(std::max(odd_up(xay, xL - yL), 2 * xL - xay), xay - 2 * yL);
    int64_t r = std::min(std::min(odd_up(xay, xU - yU), 2 * xU - xay), xay - 2 * yU);
    Lo = (int32_t)l;
    Hi = (int32_t)r;
    return l <= r;
}

bool build_band(const int64_t *a, int64_t n, int64_t lX, int64_t lY, int64_t expansion, bool dynamic, int32_t *L,
                int32_t *R) {
    return build_band(a, n, 0, 0, lX, lY, expansion, dynamic, L, R);
}

bool build_band(const int64_t *a, int64_t n, int64_t xo, int64_t yo, int64_t lX, int64_t lY, int64_t expansion,
                bool dynamic, int32_t *L, int32_t *R) {
    const int64_t D = lX + lY;
    if (!box_diag(0, 0, 0, 0, 0, L[0], R[0])) return false;
    int64_t px = 0, py = 0, e = dynamic ? 0 : expansion, xay = 1;
    for (int64_t i = 0; i <= n && xay <= D; i++) {
        int64_t nx = lX, ny = lY;
        if (i < n) {
            nx = a[3 * i] - xo + 1;
            ny = a[3 * i + 1] - yo + 1;
            if (dynamic) e = a[3 * i + 2];
        }
        const int64_t h = e / 2;
        const int64_t xL = clampz(px - h, lX), yL = clampz(ny + h, lY);
        const int64_t xU = clampz(nx + h, lX), yU = clampz(py - h, lY);
        const int64_t last = i < n ? std::min(nx + ny, D) : D;
        // (no exit inside the loop, so it vectorises: an empty anti-diagonal fails the segment)
        bool ok = true;
        for (; xay <= last; xay++) ok &= box_diag(xay, xL, yL, xU, yU, L[xay], R[xay]);
        if (!ok) return false;
        px = nx;
        py = ny;
    }
    return true;
}

void build_windows(const int32_t *L, const int32_t *R, int64_t D, const mp_band_params &p, std::vector<Window> &out) {
    out.clear();
    int64_t lo = 0;
    const int64_t narrow = 2 * p.diagonalExpansion + 1;
    int64_t xay = 1;
    while (xay <= D) {
        int64_t start = std::max(xay, lo + p.minDiagsBetweenTraceBack);
        int64_t tb = D;
        for (int64_t d = start; d < D; d++) {
            if ((R[d] - L[d]) / 2 + 1 <= narrow) {
                tb = d;
                break;
            }
        }
        if (start >= D) tb = D;
        int64_t from = tb == D ? D : tb - (p.traceBackDiagonals + 1);
        out.push_back(Window{(int32_t)tb, (int32_t)lo, (int32_t)from});
        if (tb == D) break;
        lo = from;
        xay = tb + 1;
    }
}

void split_points(const int64_t *a, int64_t n, int64_t lX, int64_t lY, int64_t maxArea, bool raggedL, bool raggedR,
                  std::vector<int64_t> &out) {
    split_points(a, n, 0, lX, lY, maxArea, raggedL, raggedR, out);
}

void split_points(const int64_t *a, int64_t n, int64_t xo, int64_t lX, int64_t lY, int64_t maxArea, bool raggedL,
                  bool raggedR, std::vector<int64_t> &out) {
    out.clear();
    const int64_t maxLen = (int64_t)std::sqrt((double)maxArea);
    int64_t x1 = 0, y1 = 0, x2 = 0, y2 = 0;
    // Returns true when the gap (x2,y2)->(x3,y3) is too large; then the block
    // ending half-way into the gap is emitted (unless skipped) and the next
    // block starts half-way back from (x3,y3).
    auto gap = [&](int64_t x3, int64_t y3, bool skip) {
        int64_t gx = x3 - x2, gy = y3 - y2;
        if (gx * gy <= maxArea) return false;
        int64_t hX = std::min(gx / 2, maxLen), hY = std::min(gy / 2, maxLen);
        if (!skip) {
            out.push_back(x1); out.push_back(y1); out.push_back(x2 + hX); out.push_back(y2 + hY);
        }
        x1 = x3 - hX;
        y1 = y3 - hY;
        return true;
    };
    for (int64_t i = 0; i < n; i++) {
        gap(a[3 * i] - xo, a[3 * i + 1], raggedL && i == 0);
        x2 = a[3 * i] - xo + 1;
        y2 = a[3 * i + 1] + 1;
    }
    if (!gap(lX, lY, raggedL && n == 0) || !raggedR) {
        out.push_back(x1); out.push_back(y1); out.push_back(lX); out.push_back(lY);
    }
}

}  // namespace mp

// ---------------------------------------------------------------------------
// C ABI: model
// ---------------------------------------------------------------------------

extern "C" mp_status mp_model_load_json(const char *path, mp_model **out) {
    if (!path || !out) {
        mp::set_error("mp_model_load_json: null argument");
        return MP_EINVAL;
    }
    *out = nullptr;
    std::ifstream fh(path, std::ios::binary);
    if (!fh) {
        mp::set_error(std::string("cannot open parameters file ") + path);
        return MP_EPARAMS;
    }
    std::stringstream ss;
    ss << fh.rdbuf();
    try {
        mpjson::Value doc = mpjson::parse(ss.str());
        if (doc.kind != mpjson::Value::Object) throw std::runtime_error("params json must be an object");
        // params_readParams2(path, requirePolish=1, requirePhase=0) (parser.c:437-528):
        // 'polish' is parsed, 'phase' skipped; without either, the document is the polish section.
        const mpjson::Value *pol = doc.get("polish");
        bool topLevel = pol || doc.get("phase");
        if (!pol && topLevel) throw std::runtime_error("did not find polish parameters in json params");
        if (topLevel)
            for (auto &kv : doc.obj)
                if (kv.first != "polish" && kv.first != "phase")
                    throw std::runtime_error("unrecognised key in params json: " + kv.first);
        mp_model *m = new mp_model();
        try {
            mp::parse_polish(pol ? *pol : doc, *m);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return MP_OK;
    } catch (const std::exception &ex) {
        mp::set_error(ex.what());
        return MP_EPARAMS;
    }
}

extern "C" mp_status mp_model_default_nucleotide(mp_model **out) {
    if (!out) {
        mp::set_error("mp_model_default_nucleotide: null argument");
        return MP_EINVAL;
    }
    mp_model *m = new mp_model();
    mp::default_transitions(m->mach[1].trans);
    mp::default_emissions(m->mach[1]);
    m->mach[1].strand = 1;
    m->mach[0] = m->mach[1];  // the default machine is used for both strands
    m->mach[0].strand = 0;
    m->repeatLog10.assign(mp::kRepeatEntries, 0.0);
    mp::default_band_params(m->p);
    *out = m;
    return MP_OK;
}

extern "C" mp_status mp_model_set_transitions(mp_model *m, const double *logTrans9, int32_t pinEnd) {
    if (!m || !logTrans9) {
        mp::set_error("mp_model_set_transitions: null argument");
        return MP_EINVAL;
    }
    for (int i = 0; i < 9; i++)
        if (std::isnan(logTrans9[i]) || logTrans9[i] > 0.0) {
            mp::set_error("mp_model_set_transitions: transitions must be log probabilities (<= 0, not NaN)");
            return MP_EINVAL;
        }
    if (m->mach[1].nStates != 3) {
        mp::set_error("mp_model_set_transitions: the model is a five-state machine (use mp_model_set_transitions5)");
        return MP_EINVAL;
    }
    if (pinEnd && !m->endPinned) std::memcpy(m->endTrans, m->mach[1].trans, sizeof(m->endTrans));
    m->endPinned = pinEnd != 0;
    for (int s = 0; s < 2; s++) std::memcpy(m->mach[s].trans, logTrans9, sizeof(m->mach[s].trans));
    return MP_OK;
}

extern "C" mp_status mp_model_set_emissions(mp_model *m, const double *logMatch16, const double *logGapX4,
                                            const double *logGapY4) {
    if (!m || !logMatch16 || !logGapX4 || !logGapY4) {
        mp::set_error("mp_model_set_emissions: null argument");
        return MP_EINVAL;
    }
    for (int i = 0; i < 24; i++) {
        const double v = i < 16 ? logMatch16[i] : (i < 20 ? logGapX4[i - 16] : logGapY4[i - 20]);
        if (std::isnan(v) || v > 0.0) {
            mp::set_error("mp_model_set_emissions: emissions must be log probabilities (<= 0, not NaN)");
            return MP_EINVAL;
        }
    }
    mp::Machine &f = m->mach[1];
    std::memcpy(f.match, logMatch16, sizeof(f.match));
    std::memcpy(f.gapX, logGapX4, sizeof(f.gapX));
    std::memcpy(f.gapY, logGapY4, sizeof(f.gapY));
    const mp::Machine r = mp::reverse_complement(f);
    std::memcpy(m->mach[0].match, r.match, sizeof(r.match));
    std::memcpy(m->mach[0].gapX, r.gapX, sizeof(r.gapX));
    std::memcpy(m->mach[0].gapY, r.gapY, sizeof(r.gapY));
    return MP_OK;
}

extern "C" mp_status mp_model_set_repeat_log10(mp_model *m, const double *log10Repeat, int32_t enable) {
    if (!m || (enable && !log10Repeat)) {
        mp::set_error("mp_model_set_repeat_log10: null argument");
        return MP_EINVAL;
    }
    if (log10Repeat) {
        for (int i = 0; i < 4 * mp::kMaxRepeat * mp::kMaxRepeat; i++)
            if (std::isnan(log10Repeat[i]) || log10Repeat[i] > 0.0) {
                mp::set_error("mp_model_set_repeat_log10: entries must be log10 probabilities (<= 0, not NaN)");
                return MP_EINVAL;
            }
        m->repeatLog10.resize(mp::kRepeatEntries, 0.0);
        std::memcpy(m->repeatLog10.data(), log10Repeat, sizeof(double) * 4 * mp::kMaxRepeat * mp::kMaxRepeat);
    }
    m->mach[0].rle = m->mach[1].rle = enable ? 1 : 0;
    return MP_OK;
}

extern "C" int32_t mp_model_num_states(const mp_model *m) { return m ? m->mach[1].nStates : 0; }

extern "C" mp_status mp_model_set_transitions5(mp_model *m, const double *logTrans25) {
    if (!m || !logTrans25) {
        mp::set_error("mp_model_set_transitions5: null argument");
        return MP_EINVAL;
    }
    for (int i = 0; i < 25; i++)
        if (std::isnan(logTrans25[i]) || logTrans25[i] > 0.0) {
            mp::set_error("mp_model_set_transitions5: transitions must be log probabilities (<= 0, not NaN)");
            return MP_EINVAL;
        }
    for (int s = 0; s < 2; s++) {
        mp::Machine &mc = m->mach[s];
        std::memcpy(mc.trans5, logTrans25, sizeof(mc.trans5));
        mc.nStates = 5;
        mp::short_gap_slots(mc);
    }
    m->endPinned = false;
    return MP_OK;
}

extern "C" mp_status mp_model_transitions5(const mp_model *m, double *logTrans25) {
    if (!m || !logTrans25) {
        mp::set_error("mp_model_transitions5: null argument");
        return MP_EINVAL;
    }
    if (m->mach[1].nStates != 5) {
        mp::set_error("mp_model_transitions5: the model is a three-state machine");
        return MP_EINVAL;
    }
    std::memcpy(logTrans25, m->mach[1].trans5, sizeof(m->mach[1].trans5));
    return MP_OK;
}

extern "C" mp_status mp_model_band_params(const mp_model *m, mp_band_params *out) {
    if (!m || !out) {
        mp::set_error("mp_model_band_params: null argument");
        return MP_EINVAL;
    }
    *out = m->p;
    return MP_OK;
}

extern "C" mp_status mp_model_tables(const mp_model *m, int32_t strand, double *trans9, double *match16,
                                     double *gapX4, double *gapY4, int32_t *rle) {
    if (!m) {
        mp::set_error("mp_model_tables: null model");
        return MP_EINVAL;
    }
    const mp::Machine &s = m->mach[strand ? 1 : 0];
    if (trans9) std::memcpy(trans9, s.trans, sizeof(s.trans));
    if (match16) std::memcpy(match16, s.match, sizeof(s.match));
    if (gapX4) std::memcpy(gapX4, s.gapX, sizeof(s.gapX));
    if (gapY4) std::memcpy(gapY4, s.gapY, sizeof(s.gapY));
    if (rle) *rle = s.rle;
    return MP_OK;
}

extern "C" double mp_model_repeat_log10(const mp_model *m, int32_t base, int32_t strand, int32_t observed,
                                        int32_t underlying) {
    if (!m || observed < 0 || underlying < 0 || observed >= mp::kMaxRepeat || underlying >= mp::kMaxRepeat ||
        base < 0 || base > 4)
        return NAN;
    int b = base == 4 ? 0 : base;  // N uses the A row (polisher.c:1758)
    int row = strand ? b : 3 - b;
    return m->repeatLog10[(size_t)row * mp::kMaxRepeat * mp::kMaxRepeat + (size_t)underlying * mp::kMaxRepeat + observed];
}

extern "C" mp_status mp_band_construct(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY,
                                       int64_t expansion, int32_t dynamic, int32_t *xmyL, int32_t *xmyR) {
    if (lX < 0 || lY < 0 || n < 0 || (n > 0 && !anchors) || !xmyL || !xmyR) {
        mp::set_error("mp_band_construct: bad argument");
        return MP_EINVAL;
    }
    if (!mp::anchors_valid(anchors, n, lX, lY)) {
        mp::set_error("mp_band_construct: invalid anchors");
        return MP_EANCHORS;
    }
    if (!mp::build_band(anchors, n, lX, lY, expansion, dynamic != 0, xmyL, xmyR)) {
        mp::set_error("mp_band_construct: invalid diagonal");
        return MP_EANCHORS;
    }
    return MP_OK;
}

extern "C" int64_t mp_split_points(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY, int64_t maxArea,
                                   int32_t raggedLeft, int32_t raggedRight, int64_t *out, int64_t cap) {
    if (lX < 0 || lY < 0 || n < 0 || (n > 0 && !anchors) || (cap > 0 && !out)) return -1;
    std::vector<int64_t> sp;
    mp::split_points(anchors, n, lX, lY, maxArea, raggedLeft != 0, raggedRight != 0, sp);
    int64_t k = (int64_t)sp.size() / 4;
    for (int64_t i = 0; i < std::min(k, cap) * 4; i++) out[i] = sp[i];
    return k;
}

extern "C" int64_t mp_traceback_windows(const int32_t *L, const int32_t *R, int64_t D, const mp_band_params *p,
                                        int32_t *out, int64_t cap) {
    if (!L || !R || !p || D < 0 || (cap > 0 && !out)) return -1;
    std::vector<mp::Window> w;
    if (D > 0) mp::build_windows(L, R, D, *p, w);
    for (int64_t i = 0; i < std::min<int64_t>((int64_t)w.size(), cap); i++) {
        out[3 * i] = w[i].tb;
        out[3 * i + 1] = w[i].lo;
        out[3 * i + 2] = w[i].from;
    }
    return (int64_t)w.size();
}

extern "C" int32_t mp_model_max_repeat_count(const mp_model *m) { return m ? mp::kMaxRepeat : 0; }

extern "C" void mp_model_free(mp_model *m) { delete m; }
