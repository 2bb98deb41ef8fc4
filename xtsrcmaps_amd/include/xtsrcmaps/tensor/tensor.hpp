// Fermi::Tensor -- the container at every boundary of the model-map path,
// with the public API subset the path uses (reference: tensor/tensor.hpp:
// 20-482): shared (shallow-copy) row-major storage, C++23 multi-index
// operator[], extent / extents / data / size / clear, begin_at / end_at and
// iterators.  Host memory, 64-byte aligned.
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdlib>
#include <functional>
#include <memory>
#include <new>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace Fermi {

namespace tensor_details {
template <typename... Ts>
concept AllIntegral = (std::is_integral_v<Ts> && ...);

template <typename T>
struct AlignedAllocator {
    using value_type = T;
    AlignedAllocator() noexcept = default;
    template <typename U>
    AlignedAllocator(const AlignedAllocator<U>&) noexcept {}
    T* allocate(std::size_t n) {
        const std::size_t bytes = ((n * sizeof(T) + 63) / 64) * 64;
        void* p = std::aligned_alloc(64, bytes ? bytes : 64);
        if (!p) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, std::size_t) noexcept { std::free(p); }
    template <typename U>
    bool operator==(const AlignedAllocator<U>&) const noexcept { return true; }
};
}  // namespace tensor_details

template <typename T, std::size_t R>
class Tensor {
  public:
    using value_type = T;
    using ExtentsType = std::array<std::size_t, R>;
    using IndicesType = std::array<long, R>;
    using AlignedVector = std::vector<T, tensor_details::AlignedAllocator<T>>;
    using iterator = typename AlignedVector::iterator;
    using const_iterator = typename AlignedVector::const_iterator;
    static constexpr std::size_t Rank = R;

    Tensor() : Tensor(ones()) {}

    explicit Tensor(ExtentsType const& extents)
        : data_(std::make_shared<AlignedVector>(count(extents))), extents_(extents) {
        strides();
    }

    template <tensor_details::AllIntegral... IndexTypes>
        requires(sizeof...(IndexTypes) == R)
    explicit Tensor(IndexTypes... extents)
        : Tensor(ExtentsType{static_cast<std::size_t>(extents)...}) {}

    Tensor(std::vector<T> const& v, ExtentsType const& extents) : Tensor(extents) {
        if (v.size() != size()) throw std::invalid_argument("Tensor: size mismatch");
        std::copy(v.begin(), v.end(), data_->begin());
    }

    template <typename U, tensor_details::AllIntegral... IndexTypes>
        requires(sizeof...(IndexTypes) == R)
    Tensor(std::vector<U> const& v, IndexTypes... extents)
        : Tensor(ExtentsType{static_cast<std::size_t>(extents)...}) {
        if (v.size() != size()) throw std::invalid_argument("Tensor: size mismatch");
        std::copy(v.begin(), v.end(), data_->begin());
    }

    template <typename InputIt>
    Tensor(InputIt first, InputIt last, ExtentsType const& extents) : Tensor(extents) {
        if ((std::size_t)std::distance(first, last) != size())
            throw std::invalid_argument("Tensor: size mismatch");
        std::copy(first, last, data_->begin());
    }

    T& operator[](IndicesType const& i) { return (*data_)[linear(i)]; }
    T const& operator[](IndicesType const& i) const { return (*data_)[linear(i)]; }

    template <tensor_details::AllIntegral... IndexTypes>
        requires(sizeof...(IndexTypes) == R)
    T& operator[](IndexTypes... i) {
        return (*data_)[linear(IndicesType{static_cast<long>(i)...})];
    }
    template <tensor_details::AllIntegral... IndexTypes>
        requires(sizeof...(IndexTypes) == R)
    T const& operator[](IndexTypes... i) const {
        return (*data_)[linear(IndicesType{static_cast<long>(i)...})];
    }

    std::size_t extent(std::size_t d) const {
        if (d >= R) throw std::out_of_range("Dimension out of range.");
        return extents_[d];
    }
    ExtentsType extents() const noexcept { return extents_; }
    ExtentsType strides_() const noexcept { return strides_v_; }
    std::size_t size() const { return count(extents_); }
    T* data() noexcept { return data_->data(); }
    T const* data() const noexcept { return data_->data(); }
    void clear() { std::fill(begin(), end(), T{}); }

    iterator begin() { return data_->begin(); }
    iterator end() { return data_->begin() + size(); }
    const_iterator begin() const { return data_->cbegin(); }
    const_iterator end() const { return data_->cbegin() + size(); }
    const_iterator cbegin() const { return begin(); }
    const_iterator cend() const { return end(); }

    // iterator at a (partial) multi-index, trailing indices zero
    template <tensor_details::AllIntegral... IndexTypes>
    iterator begin_at(IndexTypes... i) { return data_->begin() + offset(i...); }
    template <tensor_details::AllIntegral... IndexTypes>
    iterator end_at(IndexTypes... i) { return data_->begin() + offset(i...); }
    template <tensor_details::AllIntegral... IndexTypes>
    const_iterator begin_at(IndexTypes... i) const { return data_->cbegin() + offset(i...); }
    template <tensor_details::AllIntegral... IndexTypes>
    const_iterator end_at(IndexTypes... i) const { return data_->cbegin() + offset(i...); }

  private:
    static ExtentsType ones() {
        ExtentsType e;
        e.fill(1);
        return e;
    }
    static std::size_t count(ExtentsType const& e) {
        return std::accumulate(e.begin(), e.end(), std::size_t{1}, std::multiplies<>());
    }
    void strides() {
        std::size_t s = 1;
        for (std::size_t d = R; d-- > 0;) {
            strides_v_[d] = s;
            s *= extents_[d];
        }
    }
    std::size_t linear(IndicesType const& i) const {
        std::size_t k = 0;
        for (std::size_t d = 0; d < R; ++d) k += (std::size_t)i[d] * strides_v_[d];
        return k;
    }
    template <typename... I>
    std::size_t offset(I... i) const {
        static_assert(sizeof...(I) <= R, "too many indices");
        const std::size_t idx[] = {static_cast<std::size_t>(i)..., 0};
        std::size_t k = 0;
        for (std::size_t d = 0; d < sizeof...(I); ++d) k += idx[d] * strides_v_[d];
        return k;
    }

    std::shared_ptr<AlignedVector> data_;
    ExtentsType extents_{};
    ExtentsType strides_v_{};
};

}  // namespace Fermi
