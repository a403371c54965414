# ExtreMzConfig.cmake -- find_package(ExtreMz) for libExtreMz, the reference
# library's API (include/api/ExtreM.h) over the MI355X engine.  Defines the
# target name the reference exports (CMakeLists.txt:40,63-75 of the
# reference): ExtreMz::ExtreMz, a static library, so a CMake project that
# links the reference links this build unchanged:
#
#   find_package(ExtreMz REQUIRED PATHS <repo>/extrem_z_amd/cmake)
#   target_link_libraries(app PRIVATE ExtreMz::ExtreMz)
#
# The archive holds the API layer; the engine (libextremz.so) and the HIP
# runtime are its link interface.  Build both first: python __graft_entry__.py build
get_filename_component(_extremz_root "${CMAKE_CURRENT_LIST_DIR}/../.." ABSOLUTE)
set(_extremz_lib "${_extremz_root}/extrem_z_amd/_build")
if(NOT EXISTS "${_extremz_lib}/libExtreMz.a")
  set(ExtreMz_FOUND FALSE)
  set(ExtreMz_NOT_FOUND_MESSAGE "libExtreMz.a not built in ${_extremz_lib}")
  return()
endif()
if(NOT DEFINED ROCM_PATH)
  set(ROCM_PATH "/opt/rocm")
endif()
find_package(OpenMP QUIET)
if(NOT TARGET ExtreMz::ExtreMz)
  add_library(ExtreMz::ExtreMz STATIC IMPORTED)
  set(_extremz_links "${_extremz_lib}/libextremz.so" "${ROCM_PATH}/lib/libamdhip64.so")
  if(OpenMP_CXX_FOUND)
    list(APPEND _extremz_links OpenMP::OpenMP_CXX)
  endif()
  set_target_properties(ExtreMz::ExtreMz PROPERTIES
    IMPORTED_LOCATION "${_extremz_lib}/libExtreMz.a"
    INTERFACE_INCLUDE_DIRECTORIES "${_extremz_root}/include;${_extremz_root}/include/api"
    INTERFACE_LINK_LIBRARIES "${_extremz_links}"
    INTERFACE_COMPILE_FEATURES cxx_std_17)
endif()
set(ExtreMz_FOUND TRUE)
set(ExtreMz_VERSION 1.0)
