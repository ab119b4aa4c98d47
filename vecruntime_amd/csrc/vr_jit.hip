// Run-time specialization of the partitioned update (hipRTC).
//
// The generic kernels read each column's aggregate program (which state words a value updates,
// the lane's type and place in the record) from the launch arguments, so every row pays for
// uniform flag tests and for register selects of record words at run-time offsets.  Here the
// same kernel bodies (vr_agg_part.h) are instantiated with a policy whose programs and record
// layout are compile-time constants, so the compiler keeps only the operations the spec has.
// The device headers are compiled into this library as text (vr_jit_headers.inc, generated at
// build time from csrc/ and include/), so nothing is read from disk at run time.
#include <hip/hiprtc.h>

#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "vr_internal.h"
#include "vr_jit.h"
#include "vr_jit_headers.inc"

namespace vr {

namespace {

// One entry per device ordinal + source.  The map's lock is held only to find or add an entry;
// the compile runs under the entry's own lock, so a compile (seconds) holds up only the callers
// that want the same kernels, and a cached entry is returned at once.
struct JitEntry {
    std::mutex mu;
    bool done = false;
    PartJit* jit = nullptr;  // nullptr once done: the compile failed
};
std::mutex g_mu;
bool g_disabled = false;
std::map<std::string, std::shared_ptr<JitEntry>> g_cache;

std::string prog_literal(const ColProg& c) {
    char b[256];
    std::snprintf(b, sizeof(b), "{%u, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d, 0}", (unsigned)c.flags, c.col, c.type,
                  c.nulls_off, c.sum_off, c.min_off, c.max_off, c.min_bit, c.max_bit, c.rec_off, c.width);
    return b;
}

std::string jit_source(const ColProg* prog, int nprog, const RecLayout& L, bool fast_key) {
    std::string s = "#include \"vr_agg_part.h\"\nnamespace vr {\n__device__ constexpr ColProg kJitProg[PART_MAX_PROG] = {\n";
    for (int q = 0; q < PART_MAX_PROG; q++) {
        ColProg none{};
        none.nulls_off = none.sum_off = none.min_off = none.max_off = none.rec_off = -1;
        s += "    " + prog_literal(q < nprog ? prog[q] : none) + (q + 1 < PART_MAX_PROG ? ",\n" : "};\n");
    }
    s += "struct JitProgs {\n";
    s += "    static constexpr int max_rpt = PA_JIT_RPT;\n";
    s += "    static __device__ __forceinline__ int nprog(const PartArgs&) { return " + std::to_string(nprog) + "; }\n";
    s += "    static __device__ __forceinline__ const ColProg& prog(const PartArgs&, int q) { return kJitProg[q]; }\n";
    s += "    static __device__ __forceinline__ int valid_off(const PartArgs&) { return " + std::to_string(L.valid_off) +
         "; }\n";
    s += "    static __device__ __forceinline__ int row_off(const PartArgs&) { return " + std::to_string(L.row_off) +
         "; }\n";
    s += std::string("    static __device__ __forceinline__ bool fast_key(const PartArgs&) { return ") +
         (fast_key ? "true" : "false") + "; }\n";
    s += "    static __device__ __forceinline__ int rpt(const PartArgs&) { return PA_JIT_RPT; }\n";
    s += "};\n}  // namespace vr\n";
    const std::string rw = std::to_string(L.words);
    s += "template __global__ void vr::part_count_kernel<vr::JitProgs>(vr::PartArgs);\n";
    s += "template __global__ void vr::part_scatter_kernel<" + rw + ", vr::JitProgs>(vr::PartArgs);\n";
    s += "template __global__ void vr::part_merge_kernel<" + rw + ", vr::JitProgs>(vr::PartArgs);\n";
    return s;
}

std::string kernel_name(const char* kernel, const RecLayout& L) {
    return std::string("vr::") + kernel + "<" + std::to_string(L.words) + ", vr::JitProgs>";
}

PartJit* compile(int ordinal, const std::string& src, const RecLayout& L) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) return nullptr;
    hiprtcProgram prog;
    const int nh = (int)(sizeof(kJitHeaderNames) / sizeof(kJitHeaderNames[0]));
    if (hiprtcCreateProgram(&prog, src.c_str(), "vr_part_jit.hip", nh, kJitHeaderSrcs, kJitHeaderNames) !=
        HIPRTC_SUCCESS)
        return nullptr;
    const std::string nscatter = kernel_name("part_scatter_kernel", L), nmerge = kernel_name("part_merge_kernel", L);
    const std::string ncount = "vr::part_count_kernel<vr::JitProgs>";
    hiprtcAddNameExpression(prog, ncount.c_str());
    hiprtcAddNameExpression(prog, nscatter.c_str());
    hiprtcAddNameExpression(prog, nmerge.c_str());
    const std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
    // the library's own build flags (packed fp32 instructions kept out)
    const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    std::vector<char> code;
    std::string lcount, lscatter, lmerge;  // the instantiations' symbol names
    if (r == HIPRTC_SUCCESS) {
        const char *a = nullptr, *b = nullptr, *c = nullptr;
        size_t n = 0;
        if (hiprtcGetLoweredName(prog, nscatter.c_str(), &a) == HIPRTC_SUCCESS &&
            hiprtcGetLoweredName(prog, nmerge.c_str(), &b) == HIPRTC_SUCCESS &&
            hiprtcGetLoweredName(prog, ncount.c_str(), &c) == HIPRTC_SUCCESS && a && b && c &&
            hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS && n) {
            lscatter = a;
            lmerge = b;
            lcount = c;
            code.resize(n);
            if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) code.clear();
        }
    } else {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        std::fprintf(stderr, "vrgpu: run-time compile of the partitioned update failed (%s); generic kernels\n%s\n",
                     hiprtcGetErrorString(r), log.substr(0, 4000).c_str());
    }
    hiprtcDestroyProgram(&prog);
    if (code.empty()) return nullptr;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != ordinal && hipSetDevice(ordinal) != hipSuccess) return nullptr;
    PartJit* j = new PartJit{};
    bool ok = hipModuleLoadData(&j->module, code.data()) == hipSuccess &&
              hipModuleGetFunction(&j->count, j->module, lcount.c_str()) == hipSuccess &&
              hipModuleGetFunction(&j->scatter, j->module, lscatter.c_str()) == hipSuccess &&
              hipModuleGetFunction(&j->merge, j->module, lmerge.c_str()) == hipSuccess;
    if (cur != ordinal) (void)hipSetDevice(cur);
    if (!ok) {
        (void)hipGetLastError();
        delete j;
        return nullptr;
    }
    return j;
}

}  // namespace

const PartJit* part_jit(int ordinal, const ColProg* prog, int nprog, const RecLayout& L, bool fast_key) {
    const std::string src = jit_source(prog, nprog, L, fast_key);
    const std::string key = std::to_string(ordinal) + "\n" + src;
    std::shared_ptr<JitEntry> e;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (g_disabled) return nullptr;
        std::shared_ptr<JitEntry>& slot = g_cache[key];
        if (!slot) slot = std::make_shared<JitEntry>();
        e = slot;
    }
    std::lock_guard<std::mutex> g(e->mu);
    if (!e->done) {
        e->jit = compile(ordinal, src, L);
        e->done = true;
    }
    return e->jit;
}

void part_jit_disable() {
    std::lock_guard<std::mutex> g(g_mu);
    g_disabled = true;
}

}  // namespace vr
