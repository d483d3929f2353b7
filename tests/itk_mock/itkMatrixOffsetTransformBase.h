// Test-only: the mock of ITK's interface (itk_mock.h) under the name of the ITK header.
#pragma once
#include "itk_mock.h"
