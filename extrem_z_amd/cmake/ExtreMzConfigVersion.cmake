# Version 1.0, compatible with any requested 1.x or older (the reference's
# write_basic_package_version_file(... COMPATIBILITY AnyNewerVersion)).
set(PACKAGE_VERSION "1.0")
if(PACKAGE_FIND_VERSION VERSION_GREATER PACKAGE_VERSION)
  set(PACKAGE_VERSION_COMPATIBLE FALSE)
else()
  set(PACKAGE_VERSION_COMPATIBLE TRUE)
  if(PACKAGE_FIND_VERSION STREQUAL PACKAGE_VERSION)
    set(PACKAGE_VERSION_EXACT TRUE)
  endif()
endif()
