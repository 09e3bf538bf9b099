// sanitize_check.cpp -- CPU driver of the AddressSanitizer + UndefinedBehavior
// build (make -C oracle san): the oracle (oracle/vbt_oracle*.c), the host
// format library of libvbt_hip.so (vbt_amd/csrc/format.hip) and the synthetic
// encoder (vbt_amd/csrc/tools/vbt_synth.cpp), run over the files given on the
// command line, over truncated and byte-flipped copies of them, and (--encode)
// over small encodes of every field kind.  A malformed input must come back
// as an error code; the sanitizers abort on any out-of-bounds access, use of
// freed memory, leak or undefined operation.  Test infrastructure
// (tests/test_sanitize.py); prints one line per input.
#include "../../include/vbt_hip.h"
#include "../../oracle/vbt_oracle.h"
#include "../../vbt_amd/csrc/tools/vbt_synth.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

// libvbt_hip.so keeps the format library's error hook in decode.hip
static std::string g_err;
int vbt_fail(int code, const char* msg)
{
    g_err = msg ? msg : "";
    return code;
}
int vbt_fail(int code, const std::string& msg) { return vbt_fail(code, msg.c_str()); }

namespace {

struct Counts {
    uint64_t calls = 0, errors = 0;
};

// every entry point of the host format library on one file image
void format_checks(const std::vector<uint8_t>& img, Counts& n)
{
    vbt_file_header h{};
    const uint8_t none = 0;
    ++n.calls;
    if (vbt_validate_file(img.empty() ? &none : img.data(), img.size(), &h) != VBT_OK) {
        ++n.errors;
        return;
    }
    // the offsets follow the 52-byte header and the coarse AC scale table
    const uint64_t offAt = 52ull + 4ull * h.coarseAcScaleCount;
    const uint64_t payAt = offAt + 4ull * (h.leafCount + 1ull);
    const uint32_t* offs = reinterpret_cast<const uint32_t*>(img.data() + offAt);
    for (uint32_t l = 0; l < h.leafCount; ++l) {
        const uint64_t b0 = payAt + 4ull * offs[l], b1 = payAt + 4ull * offs[l + 1];
        if (b1 < b0 || b1 > img.size()) continue;
        if (h.profileType != 2) {
            vbt_leaf_view v{};
            ++n.calls;
            n.errors += vbt_parse_leaf(img.data() + b0, (size_t)(b1 - b0), 1, &v) != VBT_OK;
        }
    }
    for (uint32_t i = 0; i < 64 && h.width && h.height && h.depth; ++i)
        (void)vbt_leaf_index_for_voxel(&h, (i * 7u) % h.width, (i * 5u) % h.height, (i * 3u) % h.depth);
    if (h.width && h.height && h.depth && h.frames) {
        std::vector<uint32_t> q(4u * 500u);
        for (int pattern = 0; pattern < 4; ++pattern) {
            ++n.calls;
            if (vbt_generate_queries(&h, pattern, 500, 0x5EEDu + (uint32_t)pattern, q.data()) != VBT_OK) {
                ++n.errors;
                continue;
            }
            n.errors += vbt_sort_queries_by_leaf(&h, q.data(), 500) != VBT_OK;
        }
    }
}

// the oracle's loader, page table, caches, frames, queries and Cycles blob
void oracle_checks(const std::vector<uint8_t>& bytes, Counts& n)
{
    // the loader wants a 4-byte aligned image
    std::vector<uint32_t> words((bytes.size() + 3) / 4 + 1, 0u);
    if (!bytes.empty()) std::memcpy(words.data(), bytes.data(), bytes.size());
    const uint8_t* img = reinterpret_cast<const uint8_t*>(words.data());
    vbto_file f{};
    ++n.calls;
    if (vbto_load_file_memory(img, bytes.size(), &f) != 0) {
        ++n.errors;
        return;
    }
    const vbto_file_header& h = f.header;
    const uint32_t L = h.leafCount;
    std::vector<uint32_t> compact(L);
    const uint32_t A = vbto_page_table(&f, compact.data());
    std::vector<uint64_t> off(h.profileType == 2 ? L + 1u : A + 1u, 0u);
    for (uint32_t l = 0, r = 0; l < L; ++l) {
        const uint32_t c = vbto_leaf_cache_count(&f, l);
        if (h.profileType == 2) off[l + 1] = off[l] + c;
        else if (compact[l] != 0xFFFFFFFFu) { off[r + 1] = off[r] + c; ++r; }
    }
    const uint64_t slots = off.back();
    const uint64_t voxels = (uint64_t)h.width * h.height * h.depth;
    if (slots > (64ull << 20) || voxels > (64ull << 20)) return; // keep the checks small
    std::vector<float> cache(slots), dense(voxels), c576((uint64_t)A * 576u);
    vbto_compose comp;
    vbto_compose_default(&comp);
    const int F = (int)h.frames;
    const int frames[3] = {0, F / 2, F - 1};
    for (int fi = 0; fi < 3 && F > 0; ++fi) {
        const int t = frames[fi];
        ++n.calls;
        n.errors += vbto_frame_cache_packed(&f, t, off.data(), cache.data()) != 0;
        ++n.calls;
        n.errors += vbto_decode_frame_dense(&f, t, &comp, dense.data(), 1) != 0;
        if (h.profileType != 2) {
            ++n.calls;
            n.errors += vbto_frame_cache_576(&f, t, c576.data()) != 0;
        }
    }
    // queries inside the volume and frame range (counted), then a batch with
    // some outside them (an error by definition, not counted)
    std::mt19937 rng(7);
    std::vector<uint32_t> q(4u * 256u);
    std::vector<float> qv(256);
    for (int outside = 0; outside < 2 && h.width && h.height && h.depth && h.frames; ++outside) {
        for (uint32_t i = 0; i < 256u; ++i) {
            q[4 * i + 0] = rng() % (h.width + 2u * outside);
            q[4 * i + 1] = rng() % (h.height + 2u * outside);
            q[4 * i + 2] = rng() % (h.depth + 2u * outside);
            q[4 * i + 3] = rng() % (h.frames + 2u * outside);
        }
        const int rc = vbto_query_points(&f, q.data(), 256, &comp, qv.data(), 1);
        if (!outside) {
            ++n.calls;
            n.errors += rc != 0;
        }
    }
    if (h.profileType != 2 && F > 0) {
        uint64_t blobBytes = 0;
        ++n.calls;
        if (vbto_cycles_pack_blob(img, bytes.size(), 0, nullptr, 0, &blobBytes) == 0 && blobBytes < (256ull << 20)) {
            std::vector<uint64_t> blob((blobBytes + 7) / 8);
            if (vbto_cycles_pack_blob(img, bytes.size(), 0, blob.data(), blobBytes, &blobBytes) == 0) {
                std::vector<float> xyz(3u * 128u), out(128);
                for (uint32_t i = 0; i < 128u; ++i) {
                    xyz[3 * i + 0] = (float)(rng() % 1000u) / 1000.0f * (float)(h.width + 1u) - 0.5f;
                    xyz[3 * i + 1] = (float)(rng() % 1000u) / 1000.0f * (float)(h.height + 1u) - 0.5f;
                    xyz[3 * i + 2] = (float)(rng() % 1000u) / 1000.0f * (float)(h.depth + 1u) - 0.5f;
                }
                for (int flags = 0; flags < 4; ++flags)
                    n.errors += vbto_cycles_sample(blob.data(), blobBytes, flags, xyz.data(), 128, out.data(), 1) != 0;
            }
        } else {
            ++n.errors;
        }
    }
}

void run_all(const std::vector<uint8_t>& bytes, Counts& n)
{
    format_checks(bytes, n);
    oracle_checks(bytes, n);
}

// the file, its truncations and byte-flipped copies
bool check_file(const std::string& name, const std::vector<uint8_t>& bytes, int flips)
{
    Counts whole, bad;
    run_all(bytes, whole);
    const size_t cuts[] = {0, 8, 51, 52, 60, bytes.size() / 4, bytes.size() / 2, bytes.size() - 4, bytes.size() - 1};
    for (size_t c : cuts)
        if (c < bytes.size()) run_all(std::vector<uint8_t>(bytes.begin(), bytes.begin() + (long)c), bad);
    std::mt19937 rng(0xF11Bu);
    for (int k = 0; k < flips && bytes.size() > 64; ++k) {
        std::vector<uint8_t> v = bytes;
        for (int j = 0; j < 8; ++j) v[52 + rng() % (v.size() - 52)] ^= (uint8_t)(1u + rng() % 255u);
        run_all(v, bad);
    }
    std::printf("ok %s: %llu calls %llu errors on the file; %llu calls %llu errors on damaged copies\n",
                name.c_str(), (unsigned long long)whole.calls, (unsigned long long)whole.errors,
                (unsigned long long)bad.calls, (unsigned long long)bad.errors);
    return true;
}

bool encode_checks()
{
    bool ok = true;
    for (int kind = 0; kind <= 4; ++kind) {
        vbts_params p;
        vbts_default_params(kind, 28, 20, 18, kind == 4 ? 24 : 14, 0x5A17u + (uint64_t)kind, &p);
        void* field = vbts_field_create(&p);
        if (!field) return false;
        std::vector<float> frame((size_t)28 * 20 * 18);
        ok &= vbts_field_frame(field, 1, frame.data(), 2) == 0;
        vbts_encode_opts o;
        vbts_default_encode_opts(&o);
        o.threads = 2;
        uint8_t* bytes = nullptr;
        uint64_t size = 0;
        if (vbts_encode(field, &o, &bytes, &size) != 0) {
            std::printf("encode of kind %d failed: %s\n", kind, vbts_last_error());
            ok = false;
        } else {
            ok &= check_file("encoded kind " + std::to_string(kind), std::vector<uint8_t>(bytes, bytes + size), 8);
            vbts_free(bytes);
        }
        vbts_field_free(field);
    }
    return ok;
}

} // namespace

int main(int argc, char** argv)
{
    bool ok = true;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--encode") {
            ok &= encode_checks();
            continue;
        }
        std::ifstream in(a, std::ios::binary);
        if (!in) {
            std::printf("cannot read %s\n", a.c_str());
            return 2;
        }
        const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        ok &= check_file(a, bytes, 24);
    }
    return ok ? 0 : 1;
}
