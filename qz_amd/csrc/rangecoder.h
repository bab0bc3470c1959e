// Binary range coder with libbsc's bitstream: 32-bit range, 12-bit
// probabilities of a zero bit, output in 16-bit little-endian units with
// carry propagation, three trailing flushes; the decoder primes its code
// with three units (the first, always zero, shifts out).
//
// libbsc's coder/common/rangecoder.h is absent from the reference tree; the
// bitstream is pinned by the libbsc archive under tests/golden: decoding the
// leading 32-bit word of each of its 22 coded segments with this decoder
// yields that segment's input size from the segment table.
#pragma once
#include <cstdint>
#include <cstring>

namespace qz {

class RangeEncoder {
public:
    void init(uint8_t* out, int cap)
    {
        start_ = out;
        pos_ = out;
        end_ = out + (cap > 0 ? cap : 0);
        eob_ = out + (cap > 16 ? cap - 16 : 0);
        overflow_ = false;
        low_ = 0;
        ffnum_ = 0;
        cache_ = 0;
        range_ = 0xffffffffu;
    }
    bool eob() const { return pos_ >= eob_ || overflow_; }
    static constexpr bool kEnc = true;

    inline void bit0(int p)
    {
        range_ = (range_ >> 12) * (uint32_t)p;
        if (range_ < 0x10000u) { range_ <<= 16; shift_low(); }
    }
    inline void bit1(int p)
    {
        const uint32_t r = (range_ >> 12) * (uint32_t)p;
        low_ += r;
        range_ -= r;
        if (range_ < 0x10000u) { range_ <<= 16; shift_low(); }
    }
    // branch-free on the (unpredictable) bit: a mask selects the update
    inline int code(int bit, int p)
    {
        const uint32_t r = (range_ >> 12) * (uint32_t)p;
        const uint32_t m = 0u - (uint32_t)(bit != 0);
        low_ += r & m;
        range_ = (r & ~m) | ((range_ - r) & m);
        if (range_ < 0x10000u) { range_ <<= 16; shift_low(); }
        return bit;
    }
    void put_bit(int bit) { code(bit, 2048); }
    void put_word(uint32_t w)
    {
        for (int b = 31; b >= 0; --b) put_bit((w >> b) & 1u);
    }
    // Bytes written, or -1 if the output did not fit.
    int finish()
    {
        shift_low();
        shift_low();
        shift_low();
        return overflow_ ? -1 : (int)(pos_ - start_);
    }

private:
    inline void out16(uint32_t v)
    {
        if (pos_ + 2 <= end_) {
            pos_[0] = (uint8_t)v;
            pos_[1] = (uint8_t)(v >> 8);
        } else {
            overflow_ = true;
        }
        pos_ += 2;
    }
    void shift_low()
    {
        const uint32_t low32 = (uint32_t)low_;
        const uint32_t carry = (uint32_t)(low_ >> 32);
        if (low32 < 0xffff0000u || carry) {
            out16(cache_ + carry);
            while (ffnum_) {
                out16(carry - 1u);
                --ffnum_;
            }
            cache_ = low32 >> 16;
        } else {
            ++ffnum_;
        }
        low_ = (uint64_t)(uint32_t)(low32 << 16);
    }

    uint8_t* start_ = nullptr;
    uint8_t* pos_ = nullptr;
    uint8_t* eob_ = nullptr;
    uint8_t* end_ = nullptr;
    bool overflow_ = false;
    uint64_t low_ = 0;
    uint32_t ffnum_ = 0;
    uint32_t cache_ = 0;
    uint32_t range_ = 0;
};

class RangeDecoder {
public:
    void init(const uint8_t* in, int size)
    {
        in_ = in;
        pos_ = 0;
        size_ = size;
        range_ = 0xffffffffu;
        code_ = 0;
        for (int i = 0; i < 3; ++i) code_ = (code_ << 16) | in16();
    }
    static constexpr bool kEnc = false;
    // true once the decoder has read past the end of its input
    bool overrun() const { return pos_ > size_ + 6; }

    inline int code(int /*unused*/, int p)
    {
        const uint32_t r = (range_ >> 12) * (uint32_t)p;
        const uint32_t bit = code_ >= r ? 1u : 0u;
        const uint32_t m = 0u - bit;  // branch-free, as in the encoder
        code_ -= r & m;
        range_ = (r & ~m) | ((range_ - r) & m);
        if (range_ < 0x10000u) {
            range_ <<= 16;
            code_ = (code_ << 16) | in16();
        }
        return (int)bit;
    }
    int get_bit() { return code(0, 2048); }
    uint32_t get_word()
    {
        uint32_t w = 0;
        for (int b = 0; b < 32; ++b) w = (w << 1) | (uint32_t)get_bit();
        return w;
    }

private:
    inline uint32_t in16()
    {
        uint32_t v = 0;
        if (pos_ + 2 <= size_) v = (uint32_t)in_[pos_] | ((uint32_t)in_[pos_ + 1] << 8);
        pos_ += 2;
        return v;
    }
    const uint8_t* in_ = nullptr;
    int pos_ = 0;
    int size_ = 0;
    uint32_t range_ = 0;
    uint32_t code_ = 0;
};

}  // namespace qz
