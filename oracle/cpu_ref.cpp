// oracle/cpu_ref.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of VecRuntime's hash-aggregation and hash-join algorithm, used as the
// correctness oracle for the device library and, timed on the host, as the CPU baseline.  It
// is never linked into or called by the product library; only tests/, the package's smoke
// check and bench.py's cpu_baseline leg load it.
//
// What it restates (paths relative to the reference repository):
//   * row hashing: kernels/src/main/java/io/vecruntime/kernels/HashKernels.java:42-63,178 and
//     Decimal128.hash (Decimal128.java:155-160);
//   * GroupKeyTable: open addressing, linear probing, first-seen group ids, null key flags,
//     raw-bit double keys, rehash past 0.7 load (GroupKeyTable.java:700-764,2296-2371,2515-2561);
//     lookup without insert (:772-826); UTF8 keys as ids of a per-column dictionary filled in
//     first-use order (:1308-1430); writeKeys in group-id order (:2851-2920);
//   * GroupedAccumulators on the scatter path (more groups than GroupAssignment.LOW_CARDINALITY):
//     Count.updateAll/updateNonNull (GroupedAccumulators.java:1017-1100), LongSum with
//     Math.addExact when checked (:363-601), strict DoubleSum -- one running sum per group, rows
//     added in row order across batches (:202-206,234-236,276-280), WideLongSum's 128-bit add
//     with carry and sticky overflow (:749-781,817-824), LongMinMax / DoubleMinMax offer with
//     CompareOp.nanSafeCompare (:1248-1258,1327-1334; CompareOp.java:68-82), Decimal128MinMax
//     offer by Decimal128.compare (:1357-1462; Decimal128.java:145);
//   * the partial buffers and results the aggregate functions emit
//     (spark/src/main/scala/io/vecruntime/spark/agg/VectorAggregates.scala:118-518,905-1040);
//   * BuildTable.build's head/next chains in build order and the probe's pair order, semi/anti
//     selection and null-key rules (spark/.../sql/vecruntime/VectorHashJoinExec.scala:478-516,
//     622-633,994,1038-1182).
//   * shuffle partitioning: Spark's Pmod(Murmur3Hash(keys, seed), n) with its Murmur3_x86_32
//     routines and per-type rules, round-robin ids and the stable grouping by partition
//     (kernels/src/main/java/io/vecruntime/kernels/PartitionKernels.java:89-482).
// The decimal average's final division follows Spark's DecimalDivideWithOverflowCheck
// (org.apache.spark:spark-sql_2.13:4.1.3, not vendored in the reference): the exact quotient
// rounded half away from zero at the result scale; its double rounding through Spark's 38-digit
// intermediate cannot differ from a single rounding below 10^32 rows.
#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/vr.h"

namespace {

thread_local std::string g_error;

int fail(int status, const std::string& msg) {
    g_error = msg;
    return status;
}

// ------------------------------------------------------------------ bitmaps

inline bool bit(const uint8_t* bm, int64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }
inline void set_bit(uint8_t* bm, int64_t i, bool v) {
    if (v) bm[i >> 3] |= (uint8_t)(1u << (i & 7));
    else bm[i >> 3] &= (uint8_t)~(1u << (i & 7));
}
inline bool valid(const vr_column& c, int64_t i) { return c.validity == nullptr || bit(c.validity, i); }

// ------------------------------------------------------------------ HashKernels (HashKernels.java:42-63)

constexpr int32_t SEED = 0x2545F491;
constexpr uint32_t MULT = 0x9E3779B1u;
constexpr int32_t NULL_MARK = 0x7F4A7C15;

inline int32_t mix32(int32_t h, int32_t v) {
    uint32_t x = (uint32_t)(h ^ v) * MULT;
    return (int32_t)(x ^ (x >> 15));
}
inline int32_t finish(int32_t h) {
    uint32_t x = (uint32_t)h;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return (int32_t)x;
}
inline int32_t fold(int64_t v) { return (int32_t)((uint64_t)v ^ ((uint64_t)v >> 32)); }
// Decimal128.hash (Decimal128.java:155-160)
inline int64_t dec_hash(int64_t hi, int64_t lo) {
    uint64_t h = (uint64_t)lo * 0x9E3779B97F4A7C15ull;
    h ^= (h >> 32) ^ ((uint64_t)hi * 0xC2B2AE3D27D4EB4Full);
    h ^= h >> 29;
    return (int64_t)h;
}

// ------------------------------------------------------------------ StringDictionary

struct StringDictionary {
    std::unordered_map<std::string, int32_t> index;
    std::vector<std::string> values;
    // StringDictionary.indexOf: the id of `s`, added in first-use order when `insert`; -1 when absent.
    int32_t index_of(const std::string& s, bool insert) {
        auto it = index.find(s);
        if (it != index.end()) return it->second;
        if (!insert) return -1;
        int32_t id = (int32_t)values.size();
        index.emplace(s, id);
        values.push_back(s);
        return id;
    }
};

// The key lanes of one batch after UTF8 values are turned into table-dictionary ids.
struct KeyView {
    int32_t type;
    const uint8_t* validity;
    const void* data;
    std::vector<int32_t> ids;  // UTF8 only
    std::vector<uint8_t> idValid; // UTF8 only: 0 where the row is null (row or dictionary entry)
    bool isNull(int64_t i) const {
        if (type == VR_UTF8) return !idValid[i];
        return validity != nullptr && !bit(validity, i);
    }
    int64_t word(int64_t i) const {  // the value compared and stored (raw bits for doubles)
        switch (type) {
            case VR_INT32: return ((const int32_t*)data)[i];
            case VR_BOOL: return bit((const uint8_t*)data, i) ? 1 : 0;
            case VR_INT64:
            case VR_FLOAT64: return ((const int64_t*)data)[i];
            case VR_DECIMAL128: return ((const int64_t*)data)[2 * i];
            case VR_UTF8: return ids[i];
        }
        return 0;
    }
    int64_t word2(int64_t i) const { return type == VR_DECIMAL128 ? ((const int64_t*)data)[2 * i + 1] : 0; }
};

std::string utf8_at(const vr_column& d, int64_t e) {
    int32_t s = d.offsets[e], t = d.offsets[e + 1];
    return std::string((const char*)d.data + s, (size_t)(t - s));
}

// ------------------------------------------------------------------ GroupKeyTable

class GroupKeyTable {
   public:
    explicit GroupKeyTable(std::vector<int32_t> types) : types_(std::move(types)) {
        slots_.assign(2048, -1);  // INITIAL_CAPACITY * 2 (GroupKeyTable.java:40,597)
        mask_ = 2047;
        keys_.resize(types_.size());
        keys2_.resize(types_.size());
        nulls_.resize(types_.size());
        dicts_.resize(types_.size());
    }
    int64_t size() const { return size_; }
    const std::vector<int32_t>& types() const { return types_; }

    // toIds (GroupKeyTable.java:1308-1430): UTF8 lanes become ids of the table's dictionary.
    std::vector<KeyView> view(const vr_column* keys, int64_t n, bool insert) {
        std::vector<KeyView> v(types_.size());
        for (size_t c = 0; c < types_.size(); c++) {
            const vr_column& k = keys[c];
            v[c].type = types_[c];
            v[c].validity = k.validity;
            v[c].data = k.data;
            if (types_[c] != VR_UTF8) continue;
            v[c].ids.assign(n, 0);
            v[c].idValid.assign(n, 0);
            std::vector<int32_t> entryId;
            if (k.dictionary) entryId.assign(k.dictionary->length, INT32_MIN);
            for (int64_t i = 0; i < n; i++) {
                if (!valid(k, i)) continue;
                if (k.dictionary) {
                    int32_t e = ((const int32_t*)k.data)[i];
                    const vr_column& d = *k.dictionary;
                    if (!valid(d, e)) continue;  // a null entry: the row is null through the dictionary
                    if (entryId[e] == INT32_MIN) entryId[e] = dicts_[c].index_of(utf8_at(d, e), insert);
                    v[c].ids[i] = entryId[e];
                } else {
                    v[c].ids[i] = dicts_[c].index_of(utf8_at(k, i), insert);
                }
                v[c].idValid[i] = 1;
            }
        }
        return v;
    }

    int32_t rowHash(const std::vector<KeyView>& k, int64_t i) const {
        int32_t h = SEED;
        for (size_t c = 0; c < k.size(); c++) {
            int32_t v;
            if (k[c].isNull(i)) v = NULL_MARK;
            else {
                switch (k[c].type) {
                    case VR_INT64:
                    case VR_FLOAT64: v = fold(k[c].word(i)); break;
                    case VR_DECIMAL128: v = fold(dec_hash(k[c].word2(i), k[c].word(i))); break;
                    default: v = (int32_t)k[c].word(i); break;
                }
            }
            h = mix32(h, v);
        }
        return finish(h);
    }

    bool equals(int32_t g, const std::vector<KeyView>& k, int64_t i) const {
        for (size_t c = 0; c < k.size(); c++) {
            bool rn = k[c].isNull(i);
            if (rn != (bool)nulls_[c][g]) return false;
            if (rn) continue;
            if (keys_[c][g] != k[c].word(i)) return false;
            if (k[c].type == VR_DECIMAL128 && keys2_[c][g] != k[c].word2(i)) return false;
        }
        return true;
    }

    // lookupOrInsert (GroupKeyTable.java:2296-2323) + insert (:2515-2561)
    int32_t lookupOrInsert(const std::vector<KeyView>& k, int64_t i) {
        int32_t h = rowHash(k, i);
        size_t pos = (uint32_t)h & mask_;
        while (true) {
            int32_t g = slots_[pos];
            if (g < 0) break;
            if (hashes_[g] == h && equals(g, k, i)) return g;
            pos = (pos + 1) & mask_;
        }
        int32_t g = (int32_t)size_;
        for (size_t c = 0; c < k.size(); c++) {
            bool isNull = k[c].isNull(i);
            nulls_[c].push_back(isNull);
            keys_[c].push_back(isNull ? 0 : k[c].word(i));
            keys2_[c].push_back(isNull ? 0 : k[c].word2(i));
        }
        hashes_.push_back(h);
        slots_[pos] = g;
        size_++;
        if (size_ * 10 > (int64_t)slots_.size() * 7) rehash();
        return g;
    }

    int32_t lookupOnly(const std::vector<KeyView>& k, int64_t i) const {
        int32_t h = rowHash(k, i);
        size_t pos = (uint32_t)h & mask_;
        while (true) {
            int32_t g = slots_[pos];
            if (g < 0) return -1;
            if (hashes_[g] == h && equals(g, k, i)) return g;
            pos = (pos + 1) & mask_;
        }
    }

    // assign (GroupKeyTable.java:700-764): unselected rows get -1 and create no group.
    int64_t assign(const vr_column* keys, int64_t n, int32_t* outIds, const uint8_t* selection) {
        std::vector<KeyView> k = view(keys, n, true);
        for (int64_t i = 0; i < n; i++)
            outIds[i] = (selection && !bit(selection, i)) ? -1 : lookupOrInsert(k, i);
        return size_;
    }

    // lookup (GroupKeyTable.java:772-826): the id of the row's group or -1; returns matches.
    int64_t lookup(const vr_column* keys, int64_t n, int32_t* outIds, const uint8_t* selection) {
        std::vector<KeyView> k = view(keys, n, false);
        int64_t matched = 0;
        for (int64_t i = 0; i < n; i++) {
            bool unknown = false;
            for (size_t c = 0; c < k.size(); c++)  // a string the table never saw has id -1
                if (k[c].type == VR_UTF8 && !k[c].isNull(i) && k[c].ids[i] < 0) unknown = true;
            int32_t g = (selection && !bit(selection, i)) || unknown ? -1 : lookupOnly(k, i);
            outIds[i] = g;
            if (g >= 0) matched++;
        }
        return matched;
    }

    bool isNull(size_t c, int64_t g) const { return nulls_[c][g]; }
    int64_t key(size_t c, int64_t g) const { return keys_[c][g]; }
    int64_t key2(size_t c, int64_t g) const { return keys2_[c][g]; }
    const StringDictionary& dictionary(size_t c) const { return dicts_[c]; }

   private:
    void rehash() {  // GroupKeyTable.rehash (:2631): grow and reinsert by stored hash
        size_t cap = slots_.size() * 2;
        slots_.assign(cap, -1);
        mask_ = cap - 1;
        for (int32_t g = 0; g < size_; g++) {
            size_t pos = (uint32_t)hashes_[g] & mask_;
            while (slots_[pos] >= 0) pos = (pos + 1) & mask_;
            slots_[pos] = g;
        }
    }

    std::vector<int32_t> types_;
    std::vector<int32_t> slots_;
    size_t mask_;
    int64_t size_ = 0;
    std::vector<int32_t> hashes_;
    std::vector<std::vector<int64_t>> keys_, keys2_;
    std::vector<std::vector<uint8_t>> nulls_;
    std::vector<StringDictionary> dicts_;
};

// ------------------------------------------------------------------ 128-bit helpers

typedef __int128 i128;
typedef unsigned __int128 u128;

inline i128 make128(int64_t hi, int64_t lo) { return (i128)(((u128)(uint64_t)hi << 64) | (uint64_t)lo); }
inline int64_t hi64(i128 v) { return (int64_t)((u128)v >> 64); }
inline int64_t lo64(i128 v) { return (int64_t)(uint64_t)(u128)v; }

i128 pow10_128(int k) {
    i128 r = 1;
    for (int i = 0; i < k; i++) r *= 10;
    return r;
}

// |v| >= 10^p as a 128-bit check (WideLimbs.checked, VectorAggregates.scala:227-230)
bool past_precision(i128 v, int p) {
    if (p >= 39) return false;
    i128 lim = pow10_128(p);
    return v >= lim || v <= -lim;
}

// Unsigned 256-bit by 64-bit long division, for sum * 10^k / count at any precision.
struct U256 {
    uint64_t w[4] = {0, 0, 0, 0};  // little endian
};
U256 mul_u128_pow10(u128 m, int k) {
    U256 r;
    r.w[0] = (uint64_t)m;
    r.w[1] = (uint64_t)(m >> 64);
    for (int t = 0; t < k; t++) {
        u128 carry = 0;
        for (int j = 0; j < 4; j++) {
            u128 p = (u128)r.w[j] * 10u + carry;
            r.w[j] = (uint64_t)p;
            carry = p >> 64;
        }
    }
    return r;
}
U256 div_u256(const U256& a, uint64_t d, uint64_t* rem) {
    U256 q;
    u128 r = 0;
    for (int j = 3; j >= 0; j--) {
        u128 cur = (r << 64) | a.w[j];
        q.w[j] = (uint64_t)(cur / d);
        r = cur % d;
    }
    *rem = (uint64_t)r;
    return q;
}

// Spark's decimal average: sum (scale s) / count at scale rs, half away from zero. Returns false
// when the quotient does not fit 127 bits (then it is past every precision anyway).
bool decimal_avg(i128 sum, int64_t count, int s, int rs, i128* out) {
    bool neg = sum < 0;
    u128 mag = neg ? (u128)(-(sum + 1)) + 1 : (u128)sum;
    U256 num = mul_u128_pow10(mag, rs - s);
    uint64_t rem;
    U256 q = div_u256(num, (uint64_t)count, &rem);
    // half up on the magnitude: round away from zero when 2*rem >= count
    if ((u128)rem * 2 >= (u128)(uint64_t)count) {
        for (int j = 0; j < 4; j++)
            if (++q.w[j] != 0) break;
    }
    if (q.w[2] || q.w[3] || (q.w[1] >> 63)) return false;
    u128 qm = ((u128)q.w[1] << 64) | q.w[0];
    *out = neg ? -(i128)qm : (i128)qm;
    return true;
}

// ------------------------------------------------------------------ CompareOp.nanSafeCompare (CompareOp.java:68-82)

int nan_safe_compare(double x, double y) {
    bool xn = std::isnan(x), yn = std::isnan(y);
    if ((xn && yn) || x == y) return 0;
    if (xn) return 1;
    if (yn) return -1;
    return x > y ? 1 : -1;
}

// ------------------------------------------------------------------ accumulators (scatter path)

struct Acc {
    vr_agg_desc d;
    int32_t inType = -1;
    std::vector<int64_t> count;  // Count / LongSum count / DoubleSum count / WideLongSum count
    std::vector<int64_t> lsum;   // LongSum
    std::vector<double> dsum;    // DoubleSum
    std::vector<int64_t> lo, hi; // WideLongSum
    std::vector<uint8_t> overflowed, any, sumNull, merged;
    std::vector<int64_t> lbest;  // LongMinMax; the low limb for Decimal128MinMax
    std::vector<int64_t> hbest;  // Decimal128MinMax: the high limb
    std::vector<double> dbest;
    std::vector<int64_t> wraps;  // exact ungrouped final merge: multiples of 2^128 beyond lo/hi
    bool checkedOverflow = false;
    bool ungrouped = false;      // the handle has no grouping key

    void ensure(int64_t g) {
        if ((int64_t)count.size() >= g) return;
        count.resize(g, 0);
        lsum.resize(g, 0);
        dsum.resize(g, 0.0);
        lo.resize(g, 0);
        hi.resize(g, 0);
        overflowed.resize(g, 0);
        any.resize(g, 0);
        sumNull.resize(g, 0);
        merged.resize(g, 0);
        lbest.resize(g, 0);
        hbest.resize(g, 0);
        dbest.resize(g, 0.0);
        wraps.resize(g, 0);
    }

    // WideLongSum.add / addWide (GroupedAccumulators.java:749-824)
    void addWide(int32_t g, int64_t xh, int64_t xl) {
        int64_t l = lo[g];
        int64_t sum = (int64_t)((uint64_t)l + (uint64_t)xl);
        int64_t carry = (int64_t)(((uint64_t)((l & xl) | ((l | xl) & ~sum))) >> 63);
        int64_t h = hi[g];
        int64_t hsum = (int64_t)((uint64_t)h + (uint64_t)xh + (uint64_t)carry);
        if (((h ^ hsum) & (xh ^ hsum)) < 0) overflowed[g] = 1;
        lo[g] = sum;
        hi[g] = hsum;
    }

    // WideDecimalAvgMergeAgg.State.add (VectorAggregates.scala:580-592): a signed 128-bit overflow
    // marks the group overflowed, except in an ungrouped final, which goes on exactly (the total
    // is then wraps * 2^128 + the 128 bits).
    void mergeWide(int32_t g, int64_t xh, int64_t xl, bool exact) {
        if (!exact) {
            addWide(g, xh, xl);
            return;
        }
        int64_t h = hi[g];
        uint64_t l = (uint64_t)lo[g] + (uint64_t)xl;
        int64_t carry = l < (uint64_t)lo[g] ? 1 : 0;
        int64_t hsum = (int64_t)((uint64_t)h + (uint64_t)xh + (uint64_t)carry);
        if (((h ^ hsum) & (xh ^ hsum)) < 0) wraps[g] += h < 0 ? -1 : 1;
        lo[g] = (int64_t)l;
        hi[g] = hsum;
    }

    int64_t readLong(const vr_column& c, int64_t i) const {
        return c.type == VR_INT32 ? ((const int32_t*)c.data)[i] : ((const int64_t*)c.data)[i];
    }

    // One batch through the accumulator (the scatter loops of GroupedAccumulators).
    int update(const vr_column* values, const int32_t* ids, int64_t n) {
        const vr_column* v = d.input >= 0 ? &values[d.input] : nullptr;
        for (int64_t i = 0; i < n; i++) {
            int32_t g = ids[i];
            if (g < 0) continue;  // unselected (effectiveValidity = validity AND selection)
            switch (d.kind) {
                case VR_AGG_COUNT_STAR: count[g]++; break;
                case VR_AGG_COUNT:
                    if (valid(*v, i)) count[g]++;
                    break;
                // Merge modes (VectorAggregates.compileMerge, VectorAggregates.scala:1141-1250).
                case VR_AGG_MERGE_COUNT:  // CountMergeAgg: a LongSum over the partial counts (:1042-1062)
                    if (valid(*v, i)) count[g] = (int64_t)((uint64_t)count[g] + (uint64_t)readLong(*v, i));
                    break;
                case VR_AGG_MERGE_AVG_DOUBLE: {  // AverageMergeAgg: DoubleSum(strict) + LongSum (:1065-1090)
                    if (valid(*v, i)) dsum[g] += ((const double*)v->data)[i];
                    const vr_column& c = values[d.input2];
                    if (valid(c, i)) count[g] = (int64_t)((uint64_t)count[g] + (uint64_t)readLong(c, i));
                    break;
                }
                case VR_AGG_MERGE_SUM_WIDE: {  // WideDecimalSumMergeAgg.State.merge (:688-712)
                    const vr_column& e = values[d.input2];
                    if (!valid(e, i) || bit((const uint8_t*)e.data, i)) break;  // isEmpty null or true
                    count[g]++;                                                 // nonEmpty
                    if (!valid(*v, i)) sumNull[g] = 1;
                    else if (!sumNull[g] && !overflowed[g]) {
                        if (v->type == VR_DECIMAL128) {
                            const int64_t* p = (const int64_t*)v->data;
                            addWide(g, p[2 * i + 1], p[2 * i]);
                        } else {
                            int64_t x = readLong(*v, i);
                            addWide(g, x >> 63, x);
                        }
                    }
                    break;
                }
                case VR_AGG_MERGE_AVG_WIDE: {  // WideDecimalAvgMergeAgg.State.merge (:562-578)
                    const vr_column& c = values[d.input2];
                    if (valid(c, i)) count[g] = (int64_t)((uint64_t)count[g] + (uint64_t)readLong(c, i));
                    const bool exact = ungrouped && (d.flags & VR_AGG_RESULT);
                    if (!valid(*v, i)) sumNull[g] = 1;
                    else if (!sumNull[g] && !overflowed[g]) {
                        if (v->type == VR_DECIMAL128) {
                            const int64_t* p = (const int64_t*)v->data;
                            mergeWide(g, p[2 * i + 1], p[2 * i], exact);
                        } else {
                            int64_t x = readLong(*v, i);
                            mergeWide(g, x >> 63, x, exact);
                        }
                    }
                    break;
                }
                case VR_AGG_SUM_LONG:
                case VR_AGG_MERGE_SUM_LONG:  // SumLongAgg over the sum buffer (:1187-1188)
                    if (!valid(*v, i)) break;
                    if (d.flags & VR_AGG_CHECKED) {
                        int64_t r;
                        if (__builtin_add_overflow(lsum[g], readLong(*v, i), &r))
                            return fail(VR_ERR_ARITHMETIC_OVERFLOW, "long overflow");
                        lsum[g] = r;
                    } else {
                        lsum[g] = (int64_t)((uint64_t)lsum[g] + (uint64_t)readLong(*v, i));
                    }
                    count[g]++;
                    break;
                case VR_AGG_SUM_DOUBLE:
                case VR_AGG_MERGE_SUM_DOUBLE:  // SumDoubleAgg over the sum buffer (:1186)
                case VR_AGG_AVG_DOUBLE:
                    if (!valid(*v, i)) break;
                    dsum[g] += ((const double*)v->data)[i];  // strict: one chain, row order
                    count[g]++;
                    break;
                case VR_AGG_SUM_WIDE:
                case VR_AGG_AVG_WIDE:
                    if (!valid(*v, i)) break;
                    if (v->type == VR_DECIMAL128) {
                        const int64_t* p = (const int64_t*)v->data;
                        addWide(g, p[2 * i + 1], p[2 * i]);
                    } else {
                        int64_t x = readLong(*v, i);
                        addWide(g, x >> 63, x);
                    }
                    count[g]++;
                    break;
                case VR_AGG_MIN:
                case VR_AGG_MAX:
                case VR_AGG_MERGE_MIN:  // MinMaxAgg over the min / max buffer (:1209-1210)
                case VR_AGG_MERGE_MAX: {
                    if (!valid(*v, i)) break;
                    bool isMin = d.kind == VR_AGG_MIN || d.kind == VR_AGG_MERGE_MIN;
                    if (v->type == VR_DECIMAL128) {
                        // Decimal128MinMax.offer (GroupedAccumulators.java:1410-1422): the first
                        // value, then one strictly better by Decimal128.compare (signed high limb,
                        // then unsigned low limb, Decimal128.java:145)
                        const int64_t xl = ((const int64_t*)v->data)[2 * i], xh = ((const int64_t*)v->data)[2 * i + 1];
                        if (!any[g]) {
                            lbest[g] = xl;
                            hbest[g] = xh;
                            any[g] = 1;
                        } else {
                            const int c = xh != hbest[g] ? (xh < hbest[g] ? -1 : 1)
                                                         : ((uint64_t)xl == (uint64_t)lbest[g] ? 0 : ((uint64_t)xl < (uint64_t)lbest[g] ? -1 : 1));
                            if (isMin ? c < 0 : c > 0) {
                                lbest[g] = xl;
                                hbest[g] = xh;
                            }
                        }
                    } else if (v->type == VR_FLOAT64) {
                        double x = ((const double*)v->data)[i];
                        if (!any[g]) {
                            dbest[g] = x;
                            any[g] = 1;
                        } else {
                            int c = nan_safe_compare(x, dbest[g]);
                            if (isMin ? c < 0 : c > 0) dbest[g] = x;
                        }
                    } else {
                        int64_t x = readLong(*v, i);
                        if (!any[g]) {
                            lbest[g] = x;
                            any[g] = 1;
                        } else if (isMin ? x < lbest[g] : x > lbest[g]) {
                            lbest[g] = x;
                        }
                    }
                    break;
                }
                default: return fail(VR_ERR_UNSUPPORTED, "aggregate kind " + std::to_string(d.kind));
            }
        }
        return VR_OK;
    }
};

// Output columns an aggregate emits (Spark's partial buffer, or its result).
std::vector<int32_t> output_types(const vr_agg_desc& d, int32_t inType) {
    bool result = d.flags & VR_AGG_RESULT;
    switch (d.kind) {
        case VR_AGG_COUNT_STAR:
        case VR_AGG_COUNT:
        case VR_AGG_MERGE_COUNT: return {VR_INT64};
        case VR_AGG_SUM_LONG:
        case VR_AGG_MERGE_SUM_LONG: return {VR_INT64};
        case VR_AGG_SUM_DOUBLE:
        case VR_AGG_MERGE_SUM_DOUBLE: return {VR_FLOAT64};
        case VR_AGG_AVG_DOUBLE:
        case VR_AGG_MERGE_AVG_DOUBLE:
            if (result) return {VR_FLOAT64};
            return {VR_FLOAT64, VR_INT64};
        case VR_AGG_SUM_WIDE:
        case VR_AGG_MERGE_SUM_WIDE:
            if (result) return {VR_DECIMAL128};
            return {VR_DECIMAL128, VR_BOOL};
        case VR_AGG_AVG_WIDE:
        case VR_AGG_MERGE_AVG_WIDE:
            if (result) return {d.result_precision > 18 ? VR_DECIMAL128 : VR_INT64};
            return {VR_DECIMAL128, VR_INT64};
        case VR_AGG_MIN:
        case VR_AGG_MAX:
        case VR_AGG_MERGE_MIN:
        case VR_AGG_MERGE_MAX: return {inType};
    }
    return {};
}

void put_i64(vr_out_column& o, int64_t r, bool ok, int64_t v) {
    set_bit(o.validity, r, ok);
    ((int64_t*)o.data)[r] = ok ? v : 0;
}
void put_f64(vr_out_column& o, int64_t r, bool ok, double v) {
    set_bit(o.validity, r, ok);
    ((double*)o.data)[r] = ok ? v : 0.0;
}
void put_d128(vr_out_column& o, int64_t r, bool ok, i128 v) {
    set_bit(o.validity, r, ok);
    ((int64_t*)o.data)[2 * r] = ok ? lo64(v) : 0;
    ((int64_t*)o.data)[2 * r + 1] = ok ? hi64(v) : 0;
}

// Writes group g of one aggregate into its output columns at row r.
int emit_one(const Acc& a, int64_t g, vr_out_column* out, int64_t r) {
    const vr_agg_desc& d = a.d;
    bool result = d.flags & VR_AGG_RESULT;
    bool ansi = d.flags & VR_AGG_CHECKED;
    switch (d.kind) {
        case VR_AGG_COUNT_STAR:
        case VR_AGG_COUNT:
        case VR_AGG_MERGE_COUNT: put_i64(out[0], r, true, a.count[g]); break;  // never null
        case VR_AGG_SUM_LONG:
        case VR_AGG_MERGE_SUM_LONG: put_i64(out[0], r, a.count[g] > 0, a.lsum[g]); break;
        case VR_AGG_SUM_DOUBLE:
        case VR_AGG_MERGE_SUM_DOUBLE: put_f64(out[0], r, a.count[g] > 0, a.dsum[g]); break;
        case VR_AGG_AVG_DOUBLE:
        case VR_AGG_MERGE_AVG_DOUBLE:
            if (result) put_f64(out[0], r, a.count[g] > 0, a.dsum[g] / (double)a.count[g]);
            else {
                put_f64(out[0], r, true, a.dsum[g]);
                put_i64(out[1], r, true, a.count[g]);
            }
            break;
        case VR_AGG_MERGE_SUM_WIDE: {  // WideDecimalSumMergeAgg.State.value / write (VectorAggregates.scala:719-762)
            i128 t = make128(a.hi[g], a.lo[g]);
            const bool nonEmpty = a.count[g] > 0;
            const bool poisoned = a.sumNull[g] || a.overflowed[g];
            const bool past = past_precision(t, d.buffer_precision);
            if (result) {
                if (!nonEmpty) put_d128(out[0], r, false, 0);
                else if (poisoned) {
                    if (ansi) return fail(VR_ERR_ARITHMETIC_OVERFLOW, "overflow in sum of decimals");
                    put_d128(out[0], r, false, 0);
                } else if (past) {
                    if (ansi) return fail(VR_ERR_DECIMAL_OVERFLOW, "decimal sum past its precision");
                    put_d128(out[0], r, false, 0);
                } else put_d128(out[0], r, true, t);
            } else {
                if (!nonEmpty) put_d128(out[0], r, true, 0);  // zero while empty
                else put_d128(out[0], r, !poisoned && !past, t);
                set_bit(out[1].validity, r, true);
                set_bit((uint8_t*)out[1].data, r, !nonEmpty);  // isEmpty
            }
            break;
        }
        case VR_AGG_MERGE_AVG_WIDE: {  // WideDecimalAvgMergeAgg.State.sum / value (:596-609)
            i128 t = make128(a.hi[g], a.lo[g]);
            const bool poisoned = a.sumNull[g] || a.overflowed[g];
            const bool exact = a.ungrouped && result;  // the exact ungrouped final
            if (!result) {
                put_d128(out[0], r, !poisoned && !past_precision(t, d.buffer_precision), t);
                put_i64(out[1], r, true, a.count[g]);
                break;
            }
            bool wide = d.result_precision > 18;
            auto put_null = [&]() {
                if (wide) put_d128(out[0], r, false, 0);
                else put_i64(out[0], r, false, 0);
            };
            if (a.count[g] == 0) {  // If(count = 0, null, ...)
                put_null();
                break;
            }
            if (poisoned || (!exact && past_precision(t, d.buffer_precision))) {  // a null sum
                if (ansi) return fail(VR_ERR_ARITHMETIC_OVERFLOW, "decimal average over an overflowed sum");
                put_null();
                break;
            }
            if (exact && a.wraps[g] != 0)
                return fail(VR_ERR_UNSUPPORTED, "ungrouped decimal average over a total beyond 128 bits");
            i128 q;
            bool fits = decimal_avg(t, a.count[g], d.buffer_scale, d.result_scale, &q);
            if (!fits || past_precision(q, d.result_precision)) {
                if (ansi) return fail(VR_ERR_DECIMAL_OVERFLOW, "decimal average past its precision");
                put_null();
                break;
            }
            if (wide) put_d128(out[0], r, true, q);
            else put_i64(out[0], r, true, (int64_t)q);
            break;
        }
        case VR_AGG_SUM_WIDE: {
            i128 t = make128(a.hi[g], a.lo[g]);
            bool over = a.overflowed[g] || a.sumNull[g] || past_precision(t, d.buffer_precision);
            if (result) {
                if (a.count[g] == 0) put_d128(out[0], r, false, 0);
                else if (over) {
                    if (ansi) return fail(VR_ERR_DECIMAL_OVERFLOW, "decimal sum past its precision");
                    put_d128(out[0], r, false, 0);
                } else put_d128(out[0], r, true, t);
            } else {
                if (a.count[g] == 0) put_d128(out[0], r, true, 0);  // zero while empty
                else put_d128(out[0], r, !over, t);                  // null past the precision
                set_bit(out[1].validity, r, true);
                set_bit((uint8_t*)out[1].data, r, a.count[g] == 0);  // isEmpty
            }
            break;
        }
        case VR_AGG_AVG_WIDE: {
            i128 t = make128(a.hi[g], a.lo[g]);
            // An ungrouped result divides the total unchecked against the buffer precision, and a
            // total past 128 bits is past any precision (WideDecimalAvgAgg.sumValue,
            // VectorAggregates.scala:455-467; Escalation.total, :283-284).
            const bool ungroupedResult = a.ungrouped && result;
            bool over = a.overflowed[g] || a.sumNull[g] || (!ungroupedResult && past_precision(t, d.buffer_precision));
            if (!result) {
                if (a.count[g] == 0 && !a.sumNull[g]) put_d128(out[0], r, true, 0);
                else put_d128(out[0], r, !over, t);
                put_i64(out[1], r, true, a.count[g]);
                break;
            }
            bool wide = d.result_precision > 18;
            auto put_null = [&]() {
                if (wide) put_d128(out[0], r, false, 0);
                else put_i64(out[0], r, false, 0);
            };
            if (a.count[g] == 0) {
                put_null();
                break;
            }
            if (over) {  // the buffer could not hold the sum
                if (ansi && ungroupedResult) return fail(VR_ERR_DECIMAL_OVERFLOW, "decimal average past its precision");
                if (ansi) return fail(VR_ERR_ARITHMETIC_OVERFLOW, "decimal average over an overflowed sum");
                put_null();
                break;
            }
            i128 q;
            bool fits = decimal_avg(t, a.count[g], d.buffer_scale, d.result_scale, &q);
            if (!fits || past_precision(q, d.result_precision)) {
                if (ansi) return fail(VR_ERR_DECIMAL_OVERFLOW, "decimal average past its precision");
                put_null();
                break;
            }
            if (wide) put_d128(out[0], r, true, q);
            else put_i64(out[0], r, true, (int64_t)q);
            break;
        }
        case VR_AGG_MIN:
        case VR_AGG_MAX:
        case VR_AGG_MERGE_MIN:
        case VR_AGG_MERGE_MAX: {
            bool has = a.any[g];
            if (a.inType == VR_DECIMAL128) {
                set_bit(out[0].validity, r, has);
                ((int64_t*)out[0].data)[2 * r] = has ? a.lbest[g] : 0;
                ((int64_t*)out[0].data)[2 * r + 1] = has ? a.hbest[g] : 0;
            } else if (a.inType == VR_FLOAT64) put_f64(out[0], r, has, a.dbest[g]);
            else if (a.inType == VR_INT32) {
                set_bit(out[0].validity, r, has);
                ((int32_t*)out[0].data)[r] = has ? (int32_t)a.lbest[g] : 0;
            } else put_i64(out[0], r, has, a.lbest[g]);
            break;
        }
    }
    return VR_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

struct vro_agg {
    std::unique_ptr<GroupKeyTable> table;
    std::vector<Acc> accs;
    std::vector<int32_t> valueTypes;
    std::vector<int32_t> ids;
    // No grouping key: the ungrouped reductions of AggKernels (AggKernels.java:133-621) fold every
    // selected row into one group, and the operator emits that one row even for an empty input.
    bool ungrouped = false;
};

struct vro_join {
    std::unique_ptr<GroupKeyTable> table;
    std::vector<int32_t> head, next;
    int64_t numRows = 0;
    std::vector<int32_t> keyCols;
    std::vector<vr_column> cols;  // borrowed build columns
    bool track = false;           // buildMatched (VectorHashJoinExec.scala:1012)
    std::vector<uint8_t> matched;
};

extern "C" {

const char* vro_last_error(void) { return g_error.c_str(); }

int vro_agg_create(const vr_agg_spec* spec, vro_agg** out) {
    if (!spec || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    auto* h = new vro_agg();
    h->ungrouped = spec->num_keys == 0;
    h->table.reset(new GroupKeyTable(std::vector<int32_t>(spec->key_types, spec->key_types + spec->num_keys)));
    h->valueTypes.assign(spec->value_types, spec->value_types + spec->num_values);
    for (int a = 0; a < spec->num_aggs; a++) {
        Acc acc;
        acc.d = spec->aggs[a];
        acc.inType = acc.d.input >= 0 ? h->valueTypes[acc.d.input] : -1;
        acc.ungrouped = h->ungrouped;
        if (h->ungrouped) acc.ensure(1);
        h->accs.push_back(std::move(acc));
    }
    *out = h;
    return VR_OK;
}

// GroupKeyTable.assign + every accumulator's update for one batch; `out_ids` (nullable) receives the
// group id of each row.
int vro_agg_update(vro_agg* h, const vr_column* keys, const vr_column* values, int64_t n,
                   const uint8_t* selection, int32_t* out_ids) {
    h->ids.resize(n);
    int64_t groups = 1;
    if (h->ungrouped) {
        for (int64_t i = 0; i < n; i++) h->ids[i] = (!selection || ((selection[i >> 3] >> (i & 7)) & 1)) ? 0 : -1;
    } else {
        groups = h->table->assign(keys, n, h->ids.data(), selection);
    }
    for (auto& a : h->accs) {
        a.ensure(groups);
        int st = a.update(values, h->ids.data(), n);
        if (st != VR_OK) return st;
    }
    if (out_ids) std::memcpy(out_ids, h->ids.data(), n * sizeof(int32_t));
    return VR_OK;
}

int64_t vro_agg_num_groups(vro_agg* h) { return h->ungrouped ? 1 : h->table->size(); }

int vro_agg_output_layout(vro_agg* h, int32_t* types, int32_t max) {
    int k = 0;
    for (auto& a : h->accs)
        for (int32_t t : output_types(a.d, a.inType)) {
            if (k < max) types[k] = t;
            k++;
        }
    return k;
}

// writeKeys (GroupKeyTable.java:2851-2920) for groups [from, to) in group-id order; UTF8 keys are
// written as table-dictionary ids (INT32 lanes), then every aggregate's buffer or result.
int vro_agg_emit(vro_agg* h, int64_t from, int64_t to, vr_out_column* keys, vr_out_column* out) {
    const GroupKeyTable& t = *h->table;
    for (size_t c = 0; c < t.types().size(); c++) {
        for (int64_t g = from; g < to; g++) {
            int64_t r = g - from;
            bool ok = !t.isNull(c, g);
            set_bit(keys[c].validity, r, ok);
            switch (t.types()[c]) {
                case VR_INT32:
                case VR_UTF8: ((int32_t*)keys[c].data)[r] = (int32_t)t.key(c, g); break;
                case VR_BOOL: set_bit((uint8_t*)keys[c].data, r, t.key(c, g) != 0); break;
                case VR_INT64:
                case VR_FLOAT64: ((int64_t*)keys[c].data)[r] = t.key(c, g); break;
                case VR_DECIMAL128:
                    ((int64_t*)keys[c].data)[2 * r] = t.key(c, g);
                    ((int64_t*)keys[c].data)[2 * r + 1] = t.key2(c, g);
                    break;
            }
        }
    }
    int k = 0;
    for (auto& a : h->accs) {
        size_t w = output_types(a.d, a.inType).size();
        for (int64_t g = from; g < to; g++) {
            int st = emit_one(a, g, out + k, g - from);
            if (st != VR_OK) return st;
        }
        k += (int)w;
    }
    return VR_OK;
}

// The table dictionary of UTF8 key column c: count, then offsets/bytes when buffers are given.
int vro_agg_key_dictionary(vro_agg* h, int32_t c, int32_t* offsets, uint8_t* bytes, int64_t* count,
                           int64_t* num_bytes) {
    const StringDictionary& d = h->table->dictionary(c);
    int64_t total = 0;
    for (auto& s : d.values) total += (int64_t)s.size();
    *count = (int64_t)d.values.size();
    *num_bytes = total;
    if (offsets && bytes) {
        int64_t o = 0;
        offsets[0] = 0;
        for (size_t i = 0; i < d.values.size(); i++) {
            std::memcpy(bytes + o, d.values[i].data(), d.values[i].size());
            o += (int64_t)d.values[i].size();
            offsets[i + 1] = (int32_t)o;
        }
    }
    return VR_OK;
}

void vro_agg_destroy(vro_agg* h) { delete h; }

// BuildTable.build (VectorHashJoinExec.scala:478-516) over one concatenated build side.
int vro_join_build(const vr_join_spec* spec, const vr_column* columns, int64_t n, vro_join** out) {
    auto* j = new vro_join();
    std::vector<int32_t> kt(spec->key_types, spec->key_types + spec->num_keys);
    j->table.reset(new GroupKeyTable(kt));
    j->keyCols.assign(spec->key_columns, spec->key_columns + spec->num_keys);
    j->cols.assign(columns, columns + spec->num_columns);
    j->numRows = n;
    j->track = spec->track_matched != 0;
    j->matched.assign(n, 0);
    std::vector<vr_column> keys;
    std::vector<uint8_t> sel((n + 7) / 8, 0xFF);
    for (int32_t c : j->keyCols) keys.push_back(columns[c]);
    for (int64_t i = 0; i < n; i++)  // nonNullKeys (VectorHashJoinExec.scala:622-633)
        for (auto& k : keys)
            if (!valid(k, i)) set_bit(sel.data(), i, false);
    std::vector<int32_t> ids(n);
    int64_t groups = n > 0 ? j->table->assign(keys.data(), n, ids.data(), sel.data()) : 0;
    j->head.assign(groups, -1);
    j->next.assign(n, -1);
    for (int64_t i = n - 1; i >= 0; i--) {  // backwards so chains list rows in build order
        int32_t g = ids[i];
        if (g >= 0) {
            j->next[i] = j->head[g];
            j->head[g] = (int32_t)i;
        }
    }
    *out = j;
    return VR_OK;
}

int64_t vro_join_distinct_keys(vro_join* j) { return j->table->size(); }

// A null row of an output column (the padded side of an outer join's pair).
static void null_row(vr_out_column& dst, int64_t r) {
    set_bit(dst.validity, r, false);
    switch (dst.type) {
        case VR_INT32: ((int32_t*)dst.data)[r] = 0; break;
        case VR_BOOL: set_bit((uint8_t*)dst.data, r, false); break;
        case VR_INT64:
        case VR_FLOAT64: ((int64_t*)dst.data)[r] = 0; break;
        case VR_DECIMAL128:
            ((int64_t*)dst.data)[2 * r] = 0;
            ((int64_t*)dst.data)[2 * r + 1] = 0;
            break;
        default: break;
    }
}

static void copy_row(const vr_column& src, int64_t i, vr_out_column& dst, int64_t r) {
    set_bit(dst.validity, r, valid(src, i));
    switch (src.type) {
        case VR_INT32: ((int32_t*)dst.data)[r] = ((const int32_t*)src.data)[i]; break;
        case VR_BOOL: set_bit((uint8_t*)dst.data, r, bit((const uint8_t*)src.data, i)); break;
        case VR_INT64:
        case VR_FLOAT64: ((int64_t*)dst.data)[r] = ((const int64_t*)src.data)[i]; break;
        case VR_DECIMAL128:
            ((int64_t*)dst.data)[2 * r] = ((const int64_t*)src.data)[2 * i];
            ((int64_t*)dst.data)[2 * r + 1] = ((const int64_t*)src.data)[2 * i + 1];
            break;
        default: break;
    }
}

// VectorHashJoinIterator.probe + emitMatches / emitSemiAnti (VectorHashJoinExec.scala:1038-1182).
int vro_join_probe(vro_join* j, const vr_column* columns, int32_t num_columns, const int32_t* key_columns,
                   int64_t n, const uint8_t* selection, int32_t join_type, int32_t drop_null_keys,
                   vr_probe_output* out) {
    (void)num_columns;
    std::vector<vr_column> keys;
    for (size_t k = 0; k < j->keyCols.size(); k++) keys.push_back(columns[key_columns[k]]);
    std::vector<uint8_t> cand((n + 7) / 8, 0xFF);  // nonNullKeys: selection AND key validity
    std::vector<uint8_t> nullKey(n, 0);
    for (int64_t i = 0; i < n; i++) {
        bool ok = !selection || bit(selection, i);
        bool nk = false;
        for (auto& k : keys)
            if (!valid(k, i)) nk = true;
        nullKey[i] = nk && ok;
        set_bit(cand.data(), i, ok && !nk);
    }
    std::vector<int32_t> ids(n, -1);
    if (j->numRows > 0 && n > 0) j->table->lookup(keys.data(), n, ids.data(), cand.data());
    bool dropNull = drop_null_keys && j->numRows > 0;
    int64_t rows = 0;
    const bool pairs = join_type == VR_JOIN_INNER || join_type == VR_JOIN_OUTER;
    // PairPredicate.test (VectorHashJoinExec.scala:1287-1303): both operands non-null, each plus its
    // offset in the lane's width (wrapping), compared
    const vr_join_residual* res = out->residual;
    auto operand = [&](int32_t side, int32_t col, int64_t off, int64_t i, int64_t r, int64_t* v) -> bool {
        const vr_column& c = side == 0 ? columns[col] : j->cols[col];
        const int64_t row = side == 0 ? i : r;
        if (!valid(c, row)) return false;
        if (c.type == VR_INT32) *v = (int32_t)((uint32_t)((const int32_t*)c.data)[row] + (uint32_t)(int32_t)off);
        else *v = (int64_t)((uint64_t)((const int64_t*)c.data)[row] + (uint64_t)off);
        return true;
    };
    auto passes = [&](int64_t i, int64_t r) -> bool {
        if (!res) return true;
        int64_t x, y;
        if (!operand(res->left_side, res->left_column, res->left_offset, i, r, &x) ||
            !operand(res->right_side, res->right_column, res->right_offset, i, r, &y))
            return false;
        const int c = (x > y) - (x < y);
        switch (res->op) {
            case VR_CMP_EQ: return c == 0;
            case VR_CMP_NE: return c != 0;
            case VR_CMP_LT: return c < 0;
            case VR_CMP_LE: return c <= 0;
            case VR_CMP_GT: return c > 0;
            default: return c >= 0;
        }
    };
    auto emit = [&](int64_t i, int64_t b) {
        if (rows < out->capacity) {
            if (out->probe_idx) out->probe_idx[rows] = (int32_t)i;
            if (out->build_idx && pairs) out->build_idx[rows] = (int32_t)b;
            for (int c = 0; c < out->num_probe_out; c++) copy_row(columns[out->probe_out_columns[c]], i, out->columns[c], rows);
            for (int c = 0; c < out->num_build_out; c++) {
                vr_out_column& o = out->columns[out->num_probe_out + c];
                if (b >= 0) copy_row(j->cols[out->build_out_columns[c]], b, o, rows);
                else null_row(o, rows);
            }
        }
        rows++;
    };
    for (int64_t i = 0; i < n; i++) {
        bool sel = !selection || bit(selection, i);
        if (!sel) continue;
        if (pairs) {
            // emitMatches / the outer path (VectorHashJoinExec.scala:1160-1175): the row's build rows
            // in build order, each marked matched; a row without one is padded when preserved
            bool any = false;
            if (ids[i] >= 0)
                for (int32_t r = j->head[ids[i]]; r >= 0; r = j->next[r]) {
                    if (!passes(i, r)) continue;
                    emit(i, r);
                    any = true;
                    if (j->track) j->matched[r] = 1;
                }
            // emitOuterConditional (:1415-1450): a row none of whose pairs passes is padded
            if (!any && join_type == VR_JOIN_OUTER) emit(i, -1);
        } else {
            bool hit = false;
            if (ids[i] >= 0)
                for (int32_t r = j->head[ids[i]]; r >= 0 && !hit; r = j->next[r]) hit = passes(i, r);
            bool matched = hit || (dropNull && nullKey[i]);
            if (matched == (join_type == VR_JOIN_LEFT_SEMI)) emit(i, -1);
        }
    }
    out->rows = rows;
    if (rows > out->capacity) return fail(VR_ERR_CAPACITY, "probe output needs " + std::to_string(rows) + " rows");
    return VR_OK;
}

// emitUnmatchedBuild (VectorHashJoinExec.scala:1457-1484): (-1, r) for every build row no probe
// paired, in build order, the streamed side null.
int vro_join_unmatched_build(vro_join* j, vr_probe_output* out) {
    if (!j->track) return fail(VR_ERR_STATE, "the join does not track matched build rows");
    int64_t rows = 0;
    for (int64_t r = 0; r < j->numRows; r++) {
        if (j->matched[r]) continue;
        if (rows < out->capacity) {
            if (out->probe_idx) out->probe_idx[rows] = -1;
            if (out->build_idx) out->build_idx[rows] = (int32_t)r;
            for (int c = 0; c < out->num_probe_out; c++) null_row(out->columns[c], rows);
            for (int c = 0; c < out->num_build_out; c++)
                copy_row(j->cols[out->build_out_columns[c]], r, out->columns[out->num_probe_out + c], rows);
        }
        rows++;
    }
    out->rows = rows;
    if (rows > out->capacity) return fail(VR_ERR_CAPACITY, "unmatched build rows need " + std::to_string(rows) + " rows");
    return VR_OK;
}

void vro_join_destroy(vro_join* j) { delete j; }

// ------------------------------------------------------------------ filters

// CompareKernels.compareScalar (CompareKernels.java:57-160) turned into a selection by
// VectorExpr.selection (VectorExpr.scala:438-443): selected AND non-null AND the comparison, doubles
// in CompareOp.nanSafeCompare's order (CompareOp.java:68-82), decimals by Decimal128.compare
// (signed high limb, then unsigned low limb, Decimal128.java:145).
int vro_filter_compare(const vr_column* col, int32_t op, const vr_scalar* sc, const uint8_t* selection, int64_t n,
                       uint8_t* out, int64_t* count) {
    int64_t c = 0;
    std::memset(out, 0, (size_t)((n + 63) / 64 * 8));
    for (int64_t i = 0; i < n; i++) {
        if ((selection && !bit(selection, i)) || !valid(*col, i)) continue;
        int r = 0;
        switch (col->type) {
            case VR_INT32: {
                int32_t x = ((const int32_t*)col->data)[i], y = (int32_t)(uint32_t)(uint64_t)sc->i64;
                r = (x > y) - (x < y);
                break;
            }
            case VR_INT64: {
                int64_t x = ((const int64_t*)col->data)[i], y = sc->i64;
                r = (x > y) - (x < y);
                break;
            }
            case VR_FLOAT64: r = nan_safe_compare(((const double*)col->data)[i], sc->f64); break;
            case VR_DECIMAL128: {
                const int64_t* p = (const int64_t*)col->data + 2 * i;
                if (p[1] != sc->hi) r = p[1] > sc->hi ? 1 : -1;
                else r = ((uint64_t)p[0] > (uint64_t)sc->lo) - ((uint64_t)p[0] < (uint64_t)sc->lo);
                break;
            }
            default: return fail(VR_ERR_UNSUPPORTED, "compare lane");
        }
        bool t = false;
        switch (op) {
            case VR_CMP_EQ: t = r == 0; break;
            case VR_CMP_NE: t = r != 0; break;
            case VR_CMP_LT: t = r < 0; break;
            case VR_CMP_LE: t = r <= 0; break;
            case VR_CMP_GT: t = r > 0; break;
            case VR_CMP_GE: t = r >= 0; break;
            default: return fail(VR_ERR_INVALID_ARGUMENT, "comparison operator");
        }
        if (t) {
            set_bit(out, i, true);
            c++;
        }
    }
    if (count) *count = c;
    return VR_OK;
}

// GatherKernels.gatherFixed / gatherBits / gatherValidity / gatherUtf8 (GatherKernels.java:99-308):
// out row k = row idx[k]; -1 is a null whose fixed-width value (and BOOL bit) is zero; validity set
// unless null or padded; a UTF8 row's bytes only when valid and real.
int vro_gather(const vr_column* col, const int32_t* idx, int64_t m, vr_out_column* out) {
    int32_t pos = 0;
    if (col->type == VR_UTF8) out->offsets[0] = 0;
    for (int64_t o = 0; o < m; o++) {
        const int32_t i = idx[o];
        const bool real = i >= 0, ok = real && valid(*col, i);
        if (out->validity) set_bit(out->validity, o, ok);
        switch (col->type) {
            case VR_INT32: ((int32_t*)out->data)[o] = real ? ((const int32_t*)col->data)[i] : 0; break;
            case VR_INT64:
            case VR_FLOAT64: ((int64_t*)out->data)[o] = real ? ((const int64_t*)col->data)[i] : 0; break;
            case VR_DECIMAL128:
                ((int64_t*)out->data)[2 * o] = real ? ((const int64_t*)col->data)[2 * i] : 0;
                ((int64_t*)out->data)[2 * o + 1] = real ? ((const int64_t*)col->data)[2 * i + 1] : 0;
                break;
            case VR_BOOL: set_bit((uint8_t*)out->data, o, real && bit((const uint8_t*)col->data, i)); break;
            case VR_UTF8:
                if (ok) {
                    const int32_t a = col->offsets[i], len = col->offsets[i + 1] - a;
                    std::memcpy((uint8_t*)out->data + pos, (const uint8_t*)col->data + a, (size_t)len);
                    pos += len;
                }
                out->offsets[o + 1] = pos;
                break;
            default: return fail(VR_ERR_UNSUPPORTED, "gather lane");
        }
    }
    return VR_OK;
}

// CompactKernels.compactFixed / compactUtf8 (CompactKernels.java:108-471): the selected rows of each
// column in row order, values and validity; a dictionary-encoded column compacts its int32 codes.
int vro_compact(const vr_column* cols, int32_t ncols, int64_t n, const uint8_t* selection, vr_out_column* out,
                int64_t* rows) {
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) m += !selection || bit(selection, i);
    if (rows) *rows = m;
    for (int c = 0; c < ncols; c++) {
        const vr_column& in = cols[c];
        vr_out_column& o = out[c];
        if (o.capacity < m) return fail(VR_ERR_CAPACITY, "compacted columns need " + std::to_string(m) + " rows");
        const bool dict = in.type == VR_UTF8 && in.dictionary;
        int64_t r = 0;
        int32_t pos = 0;
        if (in.type == VR_UTF8 && !dict) o.offsets[0] = 0;
        for (int64_t i = 0; i < n; i++) {
            if (selection && !bit(selection, i)) continue;
            if (o.validity) set_bit(o.validity, r, valid(in, i));
            switch (dict ? VR_INT32 : in.type) {
                case VR_INT32: ((int32_t*)o.data)[r] = ((const int32_t*)in.data)[i]; break;
                case VR_INT64:
                case VR_FLOAT64: ((int64_t*)o.data)[r] = ((const int64_t*)in.data)[i]; break;
                case VR_DECIMAL128:
                    ((int64_t*)o.data)[2 * r] = ((const int64_t*)in.data)[2 * i];
                    ((int64_t*)o.data)[2 * r + 1] = ((const int64_t*)in.data)[2 * i + 1];
                    break;
                case VR_BOOL: set_bit((uint8_t*)o.data, r, bit((const uint8_t*)in.data, i)); break;
                default: {  // plain UTF8
                    const int32_t a = in.offsets[i], len = in.offsets[i + 1] - a;
                    std::memcpy((uint8_t*)o.data + pos, (const uint8_t*)in.data + a, (size_t)len);
                    pos += len;
                    o.offsets[r + 1] = pos;
                }
            }
            r++;
        }
    }
    return VR_OK;
}

// ------------------------------------------------------------------ CPU baseline driver

// Aggregates `num_batches` batches split round-robin over `threads` independent tables, as
// Spark's tasks each own a partial-aggregate table (VectorHashAggregateExec.scala:478,1022-1026).
// Returns the total number of groups over all tables; `seconds` receives the wall time.
int64_t vro_agg_partial_threads(const vr_agg_spec* spec, const vr_column* keys, const vr_column* values,
                                const int64_t* lengths, int32_t num_batches, int32_t threads,
                                double* seconds) {
    std::vector<int64_t> groups(threads, 0);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            vro_agg* h = nullptr;
            vro_agg_create(spec, &h);
            for (int b = t; b < num_batches; b += threads)
                vro_agg_update(h, keys + (size_t)b * spec->num_keys, values + (size_t)b * spec->num_values,
                               lengths[b], nullptr, nullptr);
            groups[t] = vro_agg_num_groups(h);
            vro_agg_destroy(h);
        });
    }
    for (auto& th : pool) th.join();
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int64_t total = 0;
    for (auto g : groups) total += g;
    return total;
}

}  // extern "C"

// ---- shuffle partitioning (PartitionKernels.java)

// Murmur3_x86_32 as Spark has it (PartitionKernels.mixK1 / mixH1 / fmix, :89-110).
static uint32_t mm_k1(uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    return k * 0x1b873593u;
}
static uint32_t mm_h1(uint32_t h, uint32_t k) {
    h ^= k;
    h = (h << 13) | (h >> 19);
    return h * 5u + 0xe6546b64u;
}
static uint32_t mm_fmix(uint32_t h, uint32_t len) {
    h ^= len;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
static uint32_t mm_int(uint32_t v, uint32_t seed) { return mm_fmix(mm_h1(seed, mm_k1(v)), 4); }  // hashInt :113-117
static uint32_t mm_long(uint64_t v, uint32_t seed) {  // hashLong :120-128: low word, then high word
    return mm_fmix(mm_h1(mm_h1(seed, mm_k1((uint32_t)v)), mm_k1((uint32_t)(v >> 32))), 8);
}
// hashUnsafeBytes (:135-150): little-endian words, then the tail a signed byte at a time.
static uint32_t mm_bytes(const uint8_t* p, int64_t len, uint32_t seed) {
    uint32_t h = seed;
    const int64_t aligned = len - len % 4;
    for (int64_t i = 0; i < aligned; i += 4) {
        uint32_t w;
        std::memcpy(&w, p + i, 4);
        h = mm_h1(h, mm_k1(w));
    }
    for (int64_t i = aligned; i < len; i++) h = mm_h1(h, mm_k1((uint32_t)(int32_t)(int8_t)p[i]));
    return mm_fmix(h, (uint32_t)len);
}

extern "C" {

// hashPartitionIds (:389-401) with mixColumn's per-kind rules (:199-250, doubleBits / floatBits
// :160-177, the decimal as BigInteger.toByteArray of the unscaled value).
int vro_hash_partition_ids(const vr_column* keys, const int32_t* kinds, int32_t nkeys, int64_t n, int32_t parts,
                           int32_t seed, int32_t* ids, int32_t* hashes) {
    if (parts < 1) return fail(VR_ERR_INVALID_ARGUMENT, "partition count below 1");
    std::vector<uint32_t> h((size_t)n, (uint32_t)seed);
    for (int c = 0; c < nkeys; c++) {
        const vr_column& k = keys[c];
        for (int64_t i = 0; i < n; i++) {
            if (!valid(k, i)) continue;  // a null key leaves the hash as it was
            uint32_t& x = h[(size_t)i];
            switch (kinds[c]) {
                case VR_HASH_INT: x = mm_int((uint32_t)((const int32_t*)k.data)[i], x); break;
                case VR_HASH_LONG: x = mm_long((uint64_t)((const int64_t*)k.data)[i], x); break;
                case VR_HASH_DOUBLE: {
                    const double v = ((const double*)k.data)[i];
                    uint64_t b;
                    std::memcpy(&b, &v, 8);
                    if (v == 0.0) b = 0;
                    else if (std::isnan(v)) b = 0x7ff8000000000000ull;
                    x = mm_long(b, x);
                    break;
                }
                case VR_HASH_FLOAT: {
                    const float f = (float)((const double*)k.data)[i];
                    uint32_t b;
                    std::memcpy(&b, &f, 4);
                    if (f == 0.0f) b = 0;
                    else if (std::isnan(f)) b = 0x7fc00000u;
                    x = mm_int(b, x);
                    break;
                }
                case VR_HASH_BOOL: x = mm_int(bit((const uint8_t*)k.data, i) ? 1 : 0, x); break;
                case VR_HASH_DECIMAL128: {
                    const uint64_t lo = ((const uint64_t*)k.data)[2 * i];
                    const int64_t hi = ((const int64_t*)k.data)[2 * i + 1];
                    const __int128 v = (__int128)(((unsigned __int128)(uint64_t)hi << 64) | lo);
                    // toByteArray: the fewest big-endian bytes whose sign bit is the value's
                    int nb = 1;
                    while (nb < 16) {
                        const __int128 lim = (__int128)1 << (8 * nb - 1);
                        if (v >= -lim && v < lim) break;
                        nb++;
                    }
                    uint8_t be[16];
                    for (int j = 0; j < nb; j++) be[j] = (uint8_t)(v >> (8 * (nb - 1 - j)));
                    x = mm_bytes(be, nb, x);
                    break;
                }
                case VR_HASH_UTF8: {
                    if (k.dictionary) {
                        const int32_t code = ((const int32_t*)k.data)[i];
                        const vr_column& dc = *k.dictionary;
                        if (code < 0 || code >= dc.length) return fail(VR_ERR_INVALID_ARGUMENT, "a dictionary code outside its dictionary");
                        x = mm_bytes((const uint8_t*)dc.data + dc.offsets[code], dc.offsets[code + 1] - dc.offsets[code], x);
                    } else {
                        x = mm_bytes((const uint8_t*)k.data + k.offsets[i], k.offsets[i + 1] - k.offsets[i], x);
                    }
                    break;
                }
                default: return fail(VR_ERR_INVALID_ARGUMENT, "unknown hash kind");
            }
        }
    }
    for (int64_t i = 0; i < n; i++) {
        if (hashes) hashes[i] = (int32_t)h[(size_t)i];
        const int32_t r = (int32_t)h[(size_t)i] % parts;  // pmod (:365-368)
        ids[i] = r < 0 ? r + parts : r;
    }
    return VR_OK;
}

// roundRobinIds (:411-422).
int vro_round_robin_ids(int64_t n, int32_t parts, int32_t start, int32_t* ids, int32_t* next) {
    int32_t p = start;
    for (int64_t i = 0; i < n; i++) {
        ids[i] = p;
        if (++p == parts) p = 0;
    }
    if (next) *next = p;
    return VR_OK;
}

// partitionOrder / partitionDestinations (:437-482): a counting sort, stable.
int vro_partition_rows(const int32_t* ids, int64_t n, int32_t parts, int32_t* rows, int32_t* dest, int64_t* starts) {
    std::vector<int64_t> cur((size_t)parts + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        if (ids[i] < 0 || ids[i] >= parts) return fail(VR_ERR_INVALID_ARGUMENT, "a partition id outside [0, num_partitions)");
        cur[(size_t)ids[i] + 1]++;
    }
    for (int32_t p = 0; p < parts; p++) cur[(size_t)p + 1] += cur[(size_t)p];
    for (int32_t p = 0; p <= parts; p++) starts[p] = cur[(size_t)p];
    for (int64_t i = 0; i < n; i++) {
        const int64_t slot = cur[(size_t)ids[i]]++;
        if (rows) rows[slot] = (int32_t)i;
        if (dest) dest[i] = (int32_t)slot;
    }
    return VR_OK;
}

}  // extern "C"
