// Roots of unity exp(-2 pi i k / M), k < M, for the register DFTs of
// fft_reg.h: cos(2 pi k/M) and -sin(2 pi k/M), correctly rounded.
// Indexed with compile-time constants only, so they fold into immediates.
#pragma once

namespace ofdft {

template <int M>
__host__ __device__ constexpr double root_re(int k);
template <int M>
__host__ __device__ constexpr double root_im(int k);

template <>
__host__ __device__ constexpr double root_re<6>(int k) {
  constexpr double t[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<6>(int k) {
  constexpr double t[6] = {0.0, -0.8660254037844386, -0.8660254037844386, 0.0, 0.8660254037844386, 0.8660254037844386};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_re<8>(int k) {
  constexpr double t[8] = {1.0, 0.7071067811865476, 0.0, -0.7071067811865476, -1.0, -0.7071067811865476, 0.0, 0.7071067811865476};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<8>(int k) {
  constexpr double t[8] = {0.0, -0.7071067811865476, -1.0, -0.7071067811865476, 0.0, 0.7071067811865476, 1.0, 0.7071067811865476};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_re<9>(int k) {
  constexpr double t[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084, -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<9>(int k) {
  constexpr double t[9] = {0.0, -0.6427876096865394, -0.984807753012208, -0.8660254037844386, -0.3420201433256687, 0.3420201433256687, 0.8660254037844386, 0.984807753012208, 0.6427876096865394};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_re<10>(int k) {
  constexpr double t[10] = {1.0, 0.8090169943749475, 0.30901699437494745, -0.30901699437494745, -0.8090169943749475, -1.0, -0.8090169943749475, -0.30901699437494745, 0.30901699437494745, 0.8090169943749475};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<10>(int k) {
  constexpr double t[10] = {0.0, -0.5877852522924731, -0.9510565162951535, -0.9510565162951535, -0.5877852522924731, 0.0, 0.5877852522924731, 0.9510565162951535, 0.9510565162951535, 0.5877852522924731};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_re<12>(int k) {
  constexpr double t[12] = {1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5, 0.0, 0.5, 0.8660254037844386};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<12>(int k) {
  constexpr double t[12] = {0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5, 0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_re<15>(int k) {
  constexpr double t[15] = {1.0, 0.9135454576426009, 0.6691306063588582, 0.30901699437494745, -0.10452846326765347, -0.5, -0.8090169943749475, -0.9781476007338057, -0.9781476007338057, -0.8090169943749475, -0.5, -0.10452846326765347, 0.30901699437494745, 0.6691306063588582, 0.9135454576426009};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<15>(int k) {
  constexpr double t[15] = {0.0, -0.4067366430758002, -0.7431448254773942, -0.9510565162951535, -0.9945218953682733, -0.8660254037844386, -0.5877852522924731, -0.20791169081775934, 0.20791169081775934, 0.5877852522924731, 0.8660254037844386, 0.9945218953682733, 0.9510565162951535, 0.7431448254773942, 0.4067366430758002};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_re<16>(int k) {
  constexpr double t[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898, 0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867};
  return t[k];
}
template <>
__host__ __device__ constexpr double root_im<16>(int k) {
  constexpr double t[16] = {0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898, 0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867, 1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898};
  return t[k];
}

}  // namespace ofdft
