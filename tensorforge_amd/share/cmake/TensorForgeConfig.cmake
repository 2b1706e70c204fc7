# find_package(TensorForge) for out-of-tree C++ builds of generated gfx950 kernels.
#
# Same package name, variables and target as the reference's config
# (tensorforge/share/cmake/TensorForgeConfig.cmake:1-50): TensorForge_INCLUDE_DIRS,
# TensorForge_SOURCES, DEVICE_BACKEND and the TensorForge::TensorForge target. Differences: the
# only backend is hip (gfx950); the runtime shim (CHECK_ERR, synchDevice, status and graph
# functions) is the prebuilt library libtensorforge_rt.so rather than a source file compiled into
# the consumer, so TensorForge_SOURCES is empty and TensorForge::TensorForge links
# TensorForge::runtime. HIP sources of a target that links TensorForge::TensorForge get the
# compile option generated kernels are built with; set CMAKE_HIP_ARCHITECTURES to gfx950
# (TensorForge_HIP_ARCHITECTURES).
include(FindPackageHandleStandardArgs)

get_filename_component(_tf_package "${CMAKE_CURRENT_LIST_DIR}/../.." REALPATH)
get_filename_component(_tf_root "${_tf_package}/.." REALPATH)

set(TensorForge_INCLUDE_DIRS "${_tf_root}/include")
set(TensorForge_SOURCES "")
set(TensorForge_HIP_ARCHITECTURES "gfx950")
set(TensorForge_LIBRARY "${_tf_package}/lib/libtensorforge_rt.so")
if(NOT EXISTS "${TensorForge_LIBRARY}")
  message(STATUS "TensorForge: ${TensorForge_LIBRARY} is not built; build it with "
                 "python -c \"from tensorforge_amd.runtime import build; build.build_runtime()\"")
  set(TensorForge_LIBRARY "TensorForge_LIBRARY-NOTFOUND")
endif()

if(NOT DEFINED DEVICE_BACKEND)
  set(DEVICE_BACKEND "hip" CACHE STRING "device backend of the generated kernels")
  set_property(CACHE DEVICE_BACKEND PROPERTY STRINGS "hip")
elseif(NOT DEVICE_BACKEND STREQUAL "hip")
  message(FATAL_ERROR "DEVICE_BACKEND=${DEVICE_BACKEND}: TensorForge here generates HIP for "
                      "gfx950 only. Allowed: hip")
endif()

find_package_handle_standard_args(TensorForge
                                  REQUIRED_VARS TensorForge_INCLUDE_DIRS TensorForge_LIBRARY)

if(TensorForge_FOUND AND NOT TARGET TensorForge::TensorForge)
  add_library(TensorForge::runtime SHARED IMPORTED)
  set_target_properties(TensorForge::runtime PROPERTIES
    IMPORTED_LOCATION "${TensorForge_LIBRARY}"
    IMPORTED_NO_SONAME TRUE)
  add_library(TensorForge::TensorForge INTERFACE IMPORTED)
  set_target_properties(TensorForge::TensorForge PROPERTIES
    INTERFACE_INCLUDE_DIRECTORIES "${TensorForge_INCLUDE_DIRS}"
    INTERFACE_LINK_LIBRARIES TensorForge::runtime
    INTERFACE_COMPILE_OPTIONS
      "$<$<COMPILE_LANGUAGE:HIP>:SHELL:-Xclang -target-feature -Xclang -packed-fp32-ops>")
endif()
