// crnd_host.h -- host-side structures of the HIP transcoder.
#pragma once
#include <stdint.h>

#include <vector>

namespace crn2 {

constexpr uint32_t kLutBits = 11;          // direct-lookup prefix width
constexpr uint32_t kLutSize = 1u << kLutBits;
constexpr uint32_t kMaxSyms = 8192;
constexpr uint32_t kMaxLevels = 16;

// A prefix-code model prepared for the device: an 11-bit lookup table whose
// entries are sym | len << 16 (0 = code longer than 11 bits or no code), plus
// the canonical tables for the long-code path.
struct DevModel {
    uint32_t lut[kLutSize];
    uint8_t len8[kLutSize];   // lut[i] >> 16: the code length alone (source of the combined prefix tables)
    uint32_t limit[17];   // limit[L] = 1 + left-aligned largest code of length L (0 if none)
    int32_t base[17];     // base[L] + code = index into sorted[]
    uint32_t num_syms;
    uint32_t pad;         // sizeof a multiple of 16: models in an array stay 16-byte aligned
};
static_assert(sizeof(DevModel) % 16 == 0, "DevModel size");

struct Header {
    uint32_t header_size, header_crc, data_size, data_crc;
    uint32_t width, height, levels, faces, format, flags, userdata0, userdata1;
    uint32_t pal_ofs[4], pal_size[4], pal_num[4];
    uint32_t tables_size, tables_ofs;
    uint32_t level_ofs[kMaxLevels];
};

// Host view of a prefix code (code sizes + canonical decode tables).
struct HostModel {
    uint32_t num_syms = 0;
    std::vector<uint8_t> sizes;
    uint32_t limit[17] = {0};
    int32_t base[17] = {0};
    std::vector<uint16_t> sorted;
    bool valid = false;
};

bool parse_header(const uint8_t* d, uint32_t size, Header& h);
bool build_model(HostModel& m, uint32_t n, const uint8_t* sizes);
void to_device_model(const HostModel& m, DevModel& dm);

// Per 11-bit prefix v (kLutSize entries): the length of the code that every
// 16-bit window starting with v starts -- 1..16, or 0 where no code matches
// (crn_decomp.h:3239-3244) -- or 0xFF where the windows differ or the model's
// canonical index check could fail for one of them (the canonical walk then
// decides).  Codes longer than the prefix change length inside a prefix only
// where a length's codes end off a 32-window boundary: at most five 0xFF
// prefixes for a model that passes the index check.
void prefix_lengths(const DevModel& dm, uint8_t* out);

// MSB-first bit reader with the reference's zero-past-the-end semantics.
struct HostBits {
    const uint8_t* next;
    const uint8_t* end;
    uint32_t buf;
    int count;
    void start(const uint8_t* p, uint32_t size)
    {
        next = p;
        end = p + size;
        buf = 0;
        count = 0;
    }
    uint32_t get_bits(uint32_t n);
    uint32_t decode_bits(uint32_t n);
    uint32_t decode(const HostModel& m);
};

bool receive_model(HostBits& b, HostModel& m);

}  // namespace crn2
