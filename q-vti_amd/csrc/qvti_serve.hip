// qvti_serve.hip -- placeholder (filled in below)
#include "qvti.h"
