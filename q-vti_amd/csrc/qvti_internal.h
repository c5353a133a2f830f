// qvti_internal.h -- declarations shared by the engine's translation units.
#pragma once

// Record a formatted message as this thread's last error and return `code`.
int qvti_fail(int code, const char* fmt, ...);
// Clear this thread's last error.
void qvti_clear_error();
