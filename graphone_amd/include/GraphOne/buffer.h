// Buffer and DenseVec: RAII arrays in host memory (arch_t::cpu) or HBM (arch_t::hip).
// Reference: include/GraphOne/base/buffer.h:12-202 and include/GraphOne/vec/dense.h:7-42 (same member
// names and meanings: data/begin/end/size/empty/set/get/to<arch>/copy_from/resize, zero-filled
// construction, element access through operator[] on the host). resize(n) keeps the reference's intent
// (reallocate and zero) without its swapped memset arguments.
#pragma once

#include <initializer_list>
#include <string>
#include <vector>

#include "GraphOne/archi.h"

namespace graph_one {

template <arch_t arch, typename value_t, typename index_t = std::size_t>
class Buffer {
public:
    Buffer() = default;

    explicit Buffer(index_t size) : size_(size), data_(archi::memalloc<arch, value_t>(size)) {
        archi::memset_zero<arch, value_t>(data_, size_);
    }

    explicit Buffer(const std::vector<value_t>& vec)
        : size_(static_cast<index_t>(vec.size())), data_(archi::memalloc<arch, value_t>(vec.size())) {
        archi::memcpy<arch, arch_t::cpu, value_t>(data_, vec.data(), size_);
    }

    Buffer(const Buffer& rhs) : size_(rhs.size_), data_(archi::memalloc<arch, value_t>(rhs.size_)) {
        archi::memcpy<arch, arch, value_t>(data_, rhs.data_, size_);
    }

    template <arch_t from>
    Buffer(const Buffer<from, value_t, index_t>& rhs)
        : size_(rhs.size()), data_(archi::memalloc<arch, value_t>(rhs.size())) {
        archi::memcpy<arch, from, value_t>(data_, rhs.data(), size_);
    }

    Buffer(Buffer&& rhs) noexcept : size_(rhs.size_), data_(rhs.data_) {
        rhs.size_ = 0;
        rhs.data_ = nullptr;
    }

    Buffer& operator=(const Buffer& rhs) {
        if (this == &rhs) return *this;
        archi::memfree<arch, value_t>(data_);
        size_ = rhs.size_;
        data_ = archi::memalloc<arch, value_t>(size_);
        archi::memcpy<arch, arch, value_t>(data_, rhs.data_, size_);
        return *this;
    }

    template <arch_t from>
    Buffer& operator=(const Buffer<from, value_t, index_t>& rhs) {
        archi::memfree<arch, value_t>(data_);
        size_ = rhs.size();
        data_ = archi::memalloc<arch, value_t>(size_);
        archi::memcpy<arch, from, value_t>(data_, rhs.data(), size_);
        return *this;
    }

    Buffer& operator=(Buffer&& rhs) noexcept {
        if (this != &rhs) {
            archi::memfree<arch, value_t>(data_);
            size_ = rhs.size_;
            data_ = rhs.data_;
            rhs.size_ = 0;
            rhs.data_ = nullptr;
        }
        return *this;
    }

    ~Buffer() { archi::memfree<arch, value_t>(data_); }

    // Copies vec (or src) into this buffer starting at element buf_begin, truncating to the buffer.
    void copy_from(const std::vector<value_t>& vec, index_t buf_begin = 0) {
        const index_t len = std::min<index_t>(static_cast<index_t>(vec.size()), size_ - buf_begin);
        archi::memcpy<arch, arch_t::cpu, value_t>(data_ + buf_begin, vec.data(), len);
    }
    void copy_from(std::initializer_list<value_t> vals, index_t buf_begin = 0) {
        copy_from(std::vector<value_t>(vals), buf_begin);
    }
    template <arch_t from>
    void copy_from(const Buffer<from, value_t, index_t>& src, index_t buf_begin = 0) {
        const index_t len = std::min<index_t>(src.size(), size_ - buf_begin);
        archi::memcpy<arch, from, value_t>(data_ + buf_begin, src.data(), len);
    }

    // Host element access: a reference for host buffers, a copied value for device buffers.
    std::conditional_t<arch == arch_t::cpu, const value_t&, value_t> operator[](index_t i) const {
        if constexpr (arch == arch_t::cpu) return data_[i];
        else return get(i);
    }
    template <arch_t a = arch>
    std::enable_if_t<a == arch_t::cpu, value_t&> operator[](index_t i) {
        return data_[i];
    }

    void set(index_t i, const value_t& v) { archi::memcpy<arch, arch_t::cpu, value_t>(data_ + i, &v, 1); }
    value_t get(index_t i) const {
        value_t v;
        archi::memcpy<arch_t::cpu, arch, value_t>(&v, data_ + i, 1);
        return v;
    }

    const value_t* begin() const { return data_; }
    value_t* begin() { return data_; }
    const value_t* end() const { return data_ + size_; }
    value_t* end() { return data_ + size_; }
    const value_t* data() const { return data_; }
    value_t* data() { return data_; }
    index_t size() const { return size_; }
    bool empty() const { return size_ == 0; }

    // Reallocates to `size` elements (when it differs) and zero-fills.
    void resize(index_t size) {
        if (size != size_) {
            archi::memfree<arch, value_t>(data_);
            data_ = archi::memalloc<arch, value_t>(size);
            size_ = size;
        }
        archi::memset_zero<arch, value_t>(data_, size_);
    }

    template <arch_t to_arch>
    Buffer<to_arch, value_t, index_t> to() const {
        return Buffer<to_arch, value_t, index_t>(*this);
    }

    std::string ToString() const {
        std::vector<value_t> h(size_);
        archi::memcpy<arch_t::cpu, arch, value_t>(h.data(), data_, size_);
        std::string s = "Buffer{ size_:" + utils::NumToString(size_) + ", data_:[";
        for (index_t i = 0; i < size_; ++i) s += (i ? "," : "") + utils::NumToString(h[i]);
        return s + "] }";
    }

protected:
    index_t size_{0};
    value_t* data_{nullptr};
};

template <arch_t arch, typename value_t, typename index_t = uint32_t>
class DenseVec : public Buffer<arch, value_t, index_t> {
public:
    using Buffer<arch, value_t, index_t>::Buffer;
    constexpr static arch_t arch_value = arch;

    DenseVec() = default;
    DenseVec(const Buffer<arch, value_t, index_t>& b) : Buffer<arch, value_t, index_t>(b) {}
    DenseVec(Buffer<arch, value_t, index_t>&& b) : Buffer<arch, value_t, index_t>(std::move(b)) {}

    template <arch_t to_arch>
    DenseVec<to_arch, value_t, index_t> to() const {
        return DenseVec<to_arch, value_t, index_t>(Buffer<to_arch, value_t, index_t>(*this));
    }

    std::string ToString() const { return "DenseVec{ arch_value:" + utils::ToString(arch) + ", " +
                                          Buffer<arch, value_t, index_t>::ToString() + " }"; }
};

}  // namespace graph_one
