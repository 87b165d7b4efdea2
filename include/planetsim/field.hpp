// Field containers with the reference's interface
// (sim/core/fields/field.hpp:54-101,168-206): bounds-checked element access
// by index or strong id, contiguous values().
#pragma once

#include <cstddef>
#include <span>
#include <type_traits>
#include <vector>

namespace planetsim {

template <typename T> class Field2D {
    static_assert(!std::is_same_v<T, bool>, "Field2D<bool> is unsupported; use an explicit byte field");

  public:
    using value_type = T;
    using size_type = std::size_t;

    Field2D() = default;
    explicit Field2D(size_type size) : values_(size) {}
    Field2D(size_type size, const T& initial) : values_(size, initial) {}
    explicit Field2D(const std::vector<T>& values) : values_(values) {}

    [[nodiscard]] bool empty() const noexcept { return values_.empty(); }
    [[nodiscard]] size_type size() const noexcept { return values_.size(); }
    [[nodiscard]] T& at(size_type i) { return values_.at(i); }
    [[nodiscard]] const T& at(size_type i) const { return values_.at(i); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] T& at(Id id) { return values_.at(id.to_index()); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] const T& at(Id id) const { return values_.at(id.to_index()); }
    [[nodiscard]] T& operator[](size_type i) { return at(i); }
    [[nodiscard]] const T& operator[](size_type i) const { return at(i); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] T& operator[](Id id) { return at(id); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] const T& operator[](Id id) const { return at(id); }
    [[nodiscard]] std::span<T> values() noexcept { return values_; }
    [[nodiscard]] std::span<const T> values() const noexcept { return values_; }

  private:
    std::vector<T> values_;
};

template <typename T> using Field = Field2D<T>;

template <typename T> class EdgeField {
    static_assert(!std::is_same_v<T, bool>, "EdgeField<bool> is unsupported; use an explicit byte field");

  public:
    using size_type = std::size_t;

    EdgeField() = default;
    explicit EdgeField(size_type size) : values_(size) {}
    EdgeField(size_type size, const T& initial) : values_(size, initial) {}

    [[nodiscard]] bool empty() const noexcept { return values_.empty(); }
    [[nodiscard]] size_type size() const noexcept { return values_.size(); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] T& at(Id id) { return values_.at(id.to_index()); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] const T& at(Id id) const { return values_.at(id.to_index()); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] T& operator[](Id id) { return at(id); }
    template <typename Id>
    requires(!std::is_integral_v<Id>) [[nodiscard]] const T& operator[](Id id) const { return at(id); }
    [[nodiscard]] std::span<T> values() noexcept { return values_; }
    [[nodiscard]] std::span<const T> values() const noexcept { return values_; }

  private:
    std::vector<T> values_;
};

}  // namespace planetsim
