#pragma once
#include <hydra/device/System.h>
